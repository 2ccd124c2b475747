// Krylov kernels for gfx950: CSR SpMV, BLAS-1 and the Jacobi-PCG of
// solver-phg.c:485 restructured into three fused kernels per iteration.
//
// Per iteration k (vector streams in brackets):
//   k_pcg_p      convergence test on the residual of iteration k-1, then
//                p = z + beta p with z = r/diag computed on the fly [r,d,p,p]
//   k_pcg_spmv   Ap = A p and per-group partial sums of (p, Ap)    [A,p,Ap]
//   k_pcg_update x += alpha p, r -= alpha Ap and per-group partial sums of
//                (r, r) and (r, r/diag)                  [x,p,r,Ap,d,x,r]
//
// Reductions: every work-group writes its partial sums with plain stores and
// the next kernel on the stream -- which needs the total anyway -- sums the
// partials in every work-group, in a fixed order, so all work-groups derive
// bit-identical scalars (alpha, beta, the residual) without device-scope
// fences or atomics.  (A release fence per work-group makes each XCD write
// back its L2, which costs more than the whole SpMV.)
//
// Solver state lives in device memory as two copies: the p-kernel of
// iteration k reads copy k&1 and writes copy (k+1)&1 from work-group 0 only,
// so no kernel reads a field another work-group of the same launch writes.
// The host launches iterations in chunks and reads the current copy from a
// pinned mailbox once per chunk; kernels of iterations after convergence
// return at once.
//
// Work-groups of the row kernels are remapped so that each of the 8 XCDs
// (own L2) processes a contiguous slice of rows, keeping the gathered x
// entries of neighbouring rows in the same L2.
#include "phg_hip_internal.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <functional>
#include <memory>
#include <type_traits>
#include <atomic>
#include <thread>
#include <vector>

namespace {

constexpr int BS = 256;      // work-group size (4 wavefronts of 64)
constexpr int NXCD = 8;
constexpr int MAX_GRID = 2048;   // partial-sum slots per reduced value
constexpr int MULTI_MAX = 16;    // matrices per batched solve

// dispatch order places work-group b on XCD b % 8; return the logical index
// that gives XCD x the x-th contiguous share of [0, gridDim.x)
__device__ inline int64_t xcd_block()
{
    int64_t nb = gridDim.x, b = blockIdx.x;
    int64_t x = b % NXCD, k = b / NXCD;
    int64_t q = nb / NXCD, rem = nb % NXCD;
    return x * q + (x < rem ? x : rem) + k;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// work-group sums of NV values, written to partials[k * gridDim.x + block]
template <int NV>
__device__ inline void write_partials(double (&v)[NV], double *partials)
{
    __shared__ double sh[NV][BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0)
            sh[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < BS / 64; w++)
            s += sh[threadIdx.x][w];
        partials[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

// totals of the np partial sums written by a previous kernel; the same fixed
// summation order in every work-group (result in every thread)
template <int NV>
__device__ inline void sum_partials(const double *partials, int np,
                                    double (&out)[NV])
{
    __shared__ double sh[NV][BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++)
        acc[k] = 0.0;
    for (int i = threadIdx.x; i < np; i += BS)
#pragma unroll
        for (int k = 0; k < NV; k++)
            acc[k] += partials[(int64_t)k * np + i];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(acc[k]);
        if (lane == 0)
            sh[k][wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < BS / 64; w++)
            s += sh[k][w];
        out[k] = s;
    }
    __syncthreads();
}

// sum_partials with f() called between the loads of the partials and their
// reduction: loads that f issues are in flight while the partials are summed,
// and since loads complete in order, waiting for the partials does not wait
// for them.  Same summation order as sum_partials (np <= MAX_GRID).
template <int NV, typename F>
__device__ inline void sum_partials_then(const double *partials, int np, double (&out)[NV],
                                         F &&f)
{
    constexpr int U = MAX_GRID / BS;
    double pv[U][NV];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int i = threadIdx.x + u * BS;
#pragma unroll
        for (int k = 0; k < NV; k++)
            pv[u][k] = i < np ? partials[(int64_t)k * np + i] : 0.0;
    }
    f();
    __shared__ double sh[NV][BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++)
        acc[k] = 0.0;
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int k = 0; k < NV; k++)
            if ((int)threadIdx.x + u * BS < np)
                acc[k] += pv[u][k];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(acc[k]);
        if (lane == 0)
            sh[k][wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < BS / 64; w++)
            s += sh[k][w];
        out[k] = s;
    }
    __syncthreads();
}

// solver state (two copies, see above)
struct State {
    double rz, rz_prev, res, tol, rb, r0;
    double rtol, btol, atol;
    int32_t nits, done, maxit, xnz;
    // reliable-update solver: largest FP32 residual since the last FP64
    // residual, update threshold, slot kind, number of FP64 residuals
    double rmax, delta;
    int32_t mode, updates;
};
// slot kinds of the reliable-update solver
constexpr int RU_START = 0, RU_ITER = 1, RU_UPDATE = 2;
constexpr int RUNNING = 0, CONVERGED = 1, BREAKDOWN = 2, ZERO_DIAG = 3;

// IsZero (utils.h:451)
__device__ inline bool is_zero(double a)
{
    return fabs(a) < 10.0 * 2.2250738585072014e-308;
}

// contiguous row range of this work-group
__device__ inline void row_range(int64_t n, int64_t &r0, int64_t &r1)
{
    int64_t b = xcd_block(), nb = gridDim.x;
    int64_t per = (n + nb - 1) / nb;
    r0 = b * per;
    r1 = r0 + per < n ? r0 + per : n;
}

// Streaming CSR SpMV over precomputed row blocks.  A work-group takes a
// contiguous range of row blocks; for each block it loads the block's
// entries (contiguous in CSR, so every load is coalesced) and stores the
// products a_ij * x_j in LDS, then one work-item per row sums its products in
// storage order -- the order of matvec.c:887-930 -- and hands (row, sum) to
// the epilogue.  A row longer than the LDS stage is a block of its own,
// summed by the whole work-group.
constexpr int NNZ_CAP = phg::SPMV_BLOCK_NNZ;   // entries staged per row block
constexpr int ROW_CAP = phg::SPMV_BLOCK_ROWS;  // rows per row block
static_assert(ROW_CAP <= BS, "one work-item per row");

// work-group barrier that orders LDS only: global loads issued before it
// (prefetches) stay in flight
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T>
__device__ inline T block_sum(T v)
{
    __shared__ T sh[BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    if (lane == 0)
        sh[wid] = v;
    lds_barrier();
    T s = 0;
#pragma unroll
    for (int w = 0; w < BS / 64; w++)
        s += sh[w];
    lds_barrier();
    return s;
}

// Sum of p[s0 .. s1) in index order (the order of matvec.c:920-928); the
// loads are issued in groups of eight so the LDS latency is paid once per
// group rather than once per entry.
template <typename T>
__device__ inline T ordered_sum(const T *p, int32_t s0, int32_t s1)
{
    T sum = 0;
    int32_t j = s0;
    for (; j + 8 <= s1; j += 8) {
        T v[8];
#pragma unroll
        for (int k = 0; k < 8; k++)
            v[k] = p[j + k];
#pragma unroll
        for (int k = 0; k < 8; k++)
            sum += v[k];
    }
    for (; j < s1; j++)
        sum += p[j];
    return sum;
}

// Row-block streaming SpMV, a family for the autotuner: CAP entries staged
// per block; MODE 0 loads them in a loop unrolled by four, MODE 1 in one
// batch of CAP / BS per work-item; in both one work-item sums a row in
// storage order (matvec.c:920-928).  MODE 2 loads like MODE 0 and sums each
// row with several lanes (different summation order, solvers only).
template <typename T, int CAP, int MODE, typename Epi>
__device__ inline void stream_spmv(const int32_t *__restrict__ rp,
                                   const int32_t *__restrict__ ci,
                                   const T *__restrict__ va,
                                   const T *__restrict__ x,
                                   const int32_t *__restrict__ rblk,
                                   int64_t nrblk, Epi epi)
{
    constexpr int PER = CAP / BS;
    constexpr int MB = 64;     // blocks whose bounds are staged in LDS at once
    static_assert(CAP % BS == 0, "whole entries per work-item");
    __shared__ T prod[CAP];
    __shared__ int32_t mrow[MB + 1], ment[MB + 1];
    int64_t b0, b1;
    row_range(nrblk, b0, b1);
    const int tid = threadIdx.x;
    for (int64_t c0 = b0; c0 < b1; c0 += MB) {
      const int nb = (int)(b1 - c0 < MB ? b1 - c0 : MB);
      lds_barrier();
      if (tid <= nb) {
          const int32_t r = rblk[c0 + tid];
          mrow[tid] = r;
          ment[tid] = rp[r];
      }
      lds_barrier();
      for (int i = 0; i < nb; i++) {
        const int32_t r0 = mrow[i], r1 = mrow[i + 1];
        const int32_t j0 = ment[i], j1 = ment[i + 1];
        if (j1 - j0 <= CAP) {
            if (MODE == 0) {
#pragma unroll 4
                for (int32_t j = j0 + tid; j < j1; j += BS)
                    prod[j - j0] = va[j] * x[ci[j]];
            }
            else {
                int32_t c[PER];
                T v[PER];
#pragma unroll
                for (int k = 0; k < PER; k++) {
                    const int32_t j = j0 + tid + BS * k;
                    c[k] = j < j1 ? ci[j] : 0;
                    v[k] = j < j1 ? va[j] : (T)0;
                }
#pragma unroll
                for (int k = 0; k < PER; k++) {
                    const int32_t j = j0 + tid + BS * k;
                    if (j < j1)
                        prod[j - j0] = v[k] * x[c[k]];
                }
            }
            lds_barrier();
            const int nrows = r1 - r0;
            if (MODE < 2) {
                if (tid < nrows) {
                    const int32_t r = r0 + tid;
                    const int32_t s0 = rp[r] - j0, s1 = rp[r + 1] - j0;
                    T sum = ordered_sum(prod, s0, s1);
                    epi(r, sum, x[r]);
                }
            }
            else {
                // lpr lanes per row (a power of two inside one wavefront),
                // strided partial sums combined with shuffles
                int lpr = 1;
                while (lpr < 16 && 2 * lpr * nrows <= BS)
                    lpr *= 2;
                const int row = tid / lpr, sub = tid % lpr;
                T sum = 0;
                int32_t r = r0 + row;
                if (row < nrows) {
                    const int32_t s0 = rp[r] - j0, s1 = rp[r + 1] - j0;
                    for (int32_t j = s0 + sub; j < s1; j += lpr)
                        sum += prod[j];
                }
                for (int o = lpr / 2; o > 0; o >>= 1)
                    sum += __shfl_xor(sum, o, 64);
                if (row < nrows && sub == 0)
                    epi(r, sum, x[r]);
            }
            lds_barrier();
        }
        else {
            // one row longer than the stage: the work-group strides over it
            T sum = 0;
            for (int32_t j = j0 + tid; j < j1; j += BS)
                sum += va[j] * x[ci[j]];
            sum = block_sum<T>(sum);
            if (tid == 0)
                epi(r0, sum, x[r0]);
        }
      }
    }
}

// Every SpMV hands (row, sum, x[row]) to its epilogue; x[row] is what the
// PCG kernels need for (p, Ap), and a layout may load it early.
// Matrix views for the SpMV-shaped kernels.  CsrView streams PHG's CSR over
// row blocks (vectors in the original numbering); SellView walks the SELL-64
// copy (vectors in the solver's permuted numbering).
template <int CAP, int MODE>
struct CsrBlockView {
    const int32_t *rp, *ci, *rblk;
    int64_t nrblk;
};
using CsrView = CsrBlockView<NNZ_CAP, 0>;

struct SellView {
    const int64_t *sptr;
    const int32_t *swidth, *scols;
    int64_t nslice, n;
};

// PHG's own CSR in storage order, walked in the solver numbering: solver
// row i is PHG row perm[i] and PHG column c is solver column iperm[c].  The
// launch shape depends on n alone, so a graph that captured it stays valid
// when a re-assembly changes the pattern (the batch path).  16 lanes per
// row, strided partial sums combined by a fixed shuffle tree.
struct PhgPermView {
    const int32_t *rp, *ci, *perm, *iperm;
    int64_t n;
};
constexpr int PPV_LANES = 16, PPV_ROWS = BS / PPV_LANES;

template <typename T, typename Epi>
__device__ inline void spmv_apply(const PhgPermView &m, const T *__restrict__ va,
                                  const T *__restrict__ x, Epi epi)
{
    const int sl = threadIdx.x % PPV_LANES;
    const int64_t step = (int64_t)gridDim.x * PPV_ROWS;
    for (int64_t i0 = (int64_t)blockIdx.x * PPV_ROWS; i0 < m.n; i0 += step) {
        const int64_t i = i0 + threadIdx.x / PPV_LANES;
        T sum = 0;
        if (i < m.n) {
            const int32_t r = m.perm[i], j1 = m.rp[r + 1];
            for (int32_t j = m.rp[r] + sl; j < j1; j += PPV_LANES)
                sum += va[j] * x[m.iperm[m.ci[j]]];
        }
        for (int o = PPV_LANES / 2; o > 0; o >>= 1)
            sum += __shfl_xor(sum, o, 64);
        if (i < m.n && sl == 0)
            epi((int32_t)i, sum, x[i]);
    }
}

template <typename T, int CAP, int MODE, typename Epi>
__device__ inline void spmv_apply(const CsrBlockView<CAP, MODE> &m,
                                  const T *va, const T *x, Epi epi)
{
    stream_spmv<T, CAP, MODE>(m.rp, m.ci, va, x, m.rblk, m.nrblk, epi);
}

// Wavefront-level CSR stream: each wavefront owns a private LDS slice and
// takes row blocks of at most WNNZ entries and 64 rows.  All loads of a block
// (8 entries per lane, coalesced, plus the lane's row bounds) are issued
// together, the products go to the wavefront's LDS slice, and lane l sums row
// r0 + l in storage order.  No work-group barrier is involved, so the four
// wavefronts of a work-group progress independently.
constexpr int WNNZ = phg::SPMV_WAVE_NNZ;   // entries per wavefront block
constexpr int WROWS = phg::SPMV_WAVE_ROWS;
static_assert(WNNZ % 64 == 0 && WROWS <= 64, "one lane per row");

struct CsrWaveView {
    const int32_t *rp, *ci, *wblk;
    int64_t nwblk;
};

template <typename T, typename Epi>
__device__ inline void spmv_apply(const CsrWaveView &m,
                                  const T *__restrict__ va,
                                  const T *__restrict__ x, Epi epi)
{
#pragma clang fp contract(off)
    __shared__ T prod_all[BS / 64][WNNZ];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    T *prod = prod_all[wv];
    const int32_t *__restrict__ rp = m.rp;
    const int32_t *__restrict__ ci = m.ci;
    int64_t g0, g1;
    row_range(m.nwblk, g0, g1);
    for (int64_t b = g0 + wv; b < g1; b += BS / 64) {
        const int32_t r0 = m.wblk[b], r1 = m.wblk[b + 1];
        const int32_t j0 = rp[r0], j1 = rp[r1];
        const int nrows = r1 - r0;
        if (j1 - j0 <= WNNZ) {
            int32_t c[WNNZ / 64];
            T v[WNNZ / 64];
            int32_t rs = 0, re = 0;
            if (lane < nrows) {
                rs = rp[r0 + lane];
                re = rp[r0 + lane + 1];
            }
#pragma unroll
            for (int k = 0; k < WNNZ / 64; k++) {
                const int32_t j = j0 + lane + 64 * k;
                c[k] = j < j1 ? ci[j] : 0;
                v[k] = j < j1 ? va[j] : (T)0;
            }
#pragma unroll
            for (int k = 0; k < WNNZ / 64; k++) {
                const int32_t j = j0 + lane + 64 * k;
                if (j < j1)
                    prod[lane + 64 * k] = v[k] * x[c[k]];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < nrows) {
                T sum = ordered_sum(prod, rs - j0, re - j0);
                epi(r0 + lane, sum, x[r0 + lane]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        else {
            // a row longer than the stage: lanes stride over it
            T sum = 0;
            for (int32_t j = j0 + lane; j < j1; j += 64)
                sum += va[j] * x[ci[j]];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1)
                sum += __shfl_xor(sum, o, 64);
            if (lane == 0)
                epi(r0, sum, x[r0]);
        }
    }
}

// SELL-64: one wavefront per slice, one row per lane; entry k of the lane's
// row sits at sptr[s] + 64 k + lane, so each load of the wavefront reads 64
// consecutive words.  The row sum runs over k in CSR storage order without
// contraction, which reproduces matvec.c:920-928 bit for bit in FP64.
template <typename T, typename Epi>
__device__ inline void spmv_apply(const SellView &m, const T *__restrict__ va,
                                  const T *__restrict__ x, Epi epi)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t g0, g1;
    row_range((m.nslice + BS / 64 - 1) / (BS / 64), g0, g1);
    for (int64_t gi = g0; gi < g1; gi++) {
        const int64_t sl = gi * (BS / 64) + wv;
        if (sl >= m.nslice)
            break;
        const int64_t base = m.sptr[sl] + lane;
        const int w = m.swidth[sl];
        const int32_t *__restrict__ cp = m.scols + base;
        const T *__restrict__ vp = va + base;
        T acc = 0;
        int k = 0;
        for (; k + 4 <= w; k += 4) {
            const int32_t c0 = cp[0], c1 = cp[64], c2 = cp[128], c3 = cp[192];
            const T v0 = vp[0], v1 = vp[64], v2 = vp[128], v3 = vp[192];
            const T x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
            acc += v0 * x0;
            acc += v1 * x1;
            acc += v2 * x2;
            acc += v3 * x3;
            cp += 256;
            vp += 256;
        }
        for (; k < w; k++) {
            acc += vp[0] * x[cp[0]];
            cp += 64;
            vp += 64;
        }
        const int64_t i = sl * 64 + lane;
        if (i < m.n)
            epi((int32_t)i, acc, x[i]);
    }
}

// Dictionary-compressed CSR stream (layout built in dictcsr.hip).  A
// work-group takes a contiguous range of blocks.  Per block each work-item
// loads 8 consecutive 16-bit indices and 8 values with 16-byte loads (two
// groups of 8 per work-item for a 4096-entry block); the layout stores each
// wavefront's 512 entries transposed, so the j-th gather of a wavefront
// reads 64 consecutive entries of the block (few cache lines).  Columns are
// decoded through the block's dictionary in LDS, the products go to LDS in
// storage order, and one work-item per row sums its products in order.  The
// loads of the next block are issued before the products of the current
// one are formed, so the matrix stream stays in flight across the LDS
// barriers and the row sums.
struct DictView {
    const int32_t *rp, *brow, *bent, *bdict, *dict;
    const uint16_t *idx;
    int64_t nblk;
};

template <typename T> struct Vals8 { T v[8]; };

template <typename T>
__device__ inline Vals8<T> load8(const T *p)
{
    Vals8<T> r;
    if (sizeof(T) == 4) {
        const float4 a = ((const float4 *)p)[0], b = ((const float4 *)p)[1];
        r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
        r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
    }
    else {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double2 a = ((const double2 *)p)[k];
            r.v[2 * k] = a.x;
            r.v[2 * k + 1] = a.y;
        }
    }
    return r;
}

template <typename T, int CAP = phg::DICT_BLOCK_NNZ> struct DictStage {
    static constexpr int NCH = CAP / (BS * 8);
    static_assert(NCH >= 1 && NCH * BS * 8 == CAP, "block capacity");
    uint4 ib[NCH];
    Vals8<T> vb[NCH];
    int32_t d, f0, f1, q0, q1, s0, s1;
    T xr;          // x of the work-item's row (for the epilogue)
};

// blocks whose metadata a work-group stages in LDS at a time
constexpr int DICT_META = 64;

// loads block (entries [f0, f1), rows [q0, q1), dictionary slots [d0, d1))
// into registers: 8 indices and 8 values per work-item and chunk, the
// work-item's dictionary slot, and the row pointers and x of its row
template <typename T, int CAP>
__device__ inline void stage_fetch(const DictView &m, const T *__restrict__ va,
                                   const T *__restrict__ x, int32_t f0,
                                   int32_t f1, int32_t q0, int32_t q1,
                                   int32_t d0, int32_t d1,
                                   DictStage<T, CAP> &st)
{
    const int tid = threadIdx.x;
    st.f0 = f0;
    st.f1 = f1;
    st.q0 = q0;
    st.q1 = q1;
#pragma unroll
    for (int k = 0; k < DictStage<T, CAP>::NCH; k++) {
        const int32_t e = st.f0 + (tid + BS * k) * 8;
        if (e < st.f1) {
            st.ib[k] = *(const uint4 *)(m.idx + e);
            st.vb[k] = load8(va + e);
        }
    }
    const int32_t nd = d1 - d0;
    st.d = tid < nd ? m.dict[d0 + tid] : 0;
    const int32_t r = st.q0 + tid;
    if (r < st.q1) {
        st.s0 = m.rp[r];
        st.s1 = m.rp[r + 1];
        st.xr = x[r];
    }
}

// block metadata of up to DICT_META blocks of a sweep, staged in LDS
struct SweepMeta {
    int32_t f0[DICT_META], f1[DICT_META], q0[DICT_META], q1[DICT_META];
    int32_t d0[DICT_META], d1[DICT_META];
};

// stages the metadata of blocks c0, c0 + step, ... (at most DICT_META, all
// before end) and returns their number; brackets the writes with LDS
// barriers so the previous chunk is done with md
__device__ inline int sweep_meta(const DictView &m, int64_t c0, int64_t end,
                                 int64_t step, SweepMeta &md)
{
    static_assert(DICT_META <= BS, "one work-item per metadata entry");
    const int64_t left = (end - c0 + step - 1) / step;
    const int nb = (int)(left < DICT_META ? left : DICT_META);
    const int tid = threadIdx.x;
    lds_barrier();
    if (tid < nb) {
        const int64_t blk = c0 + (int64_t)tid * step;
        md.f0[tid] = m.bent[blk];
        md.f1[tid] = m.bent[blk + 1];
        md.q0[tid] = m.brow[blk];
        md.q1[tid] = m.brow[blk + 1];
        md.d0[tid] = m.bdict[blk];
        md.d1[tid] = m.bdict[blk + 1];
    }
    lds_barrier();
    return nb;
}

template <typename T, int CAP>
__device__ inline void sweep_fetch(const DictView &m, const SweepMeta &md,
                                   const T *__restrict__ va,
                                   const T *__restrict__ x, int i,
                                   DictStage<T, CAP> &st)
{
    stage_fetch(m, va, x, md.f0[i], md.f1[i], md.q0[i], md.q1[i], md.d0[i],
                md.d1[i], st);
}

template <typename T>
__device__ inline void store8(T *p, const T (&v)[8])
{
    if (sizeof(T) == 4) {
        ((float4 *)p)[0] = make_float4(v[0], v[1], v[2], v[3]);
        ((float4 *)p)[1] = make_float4(v[4], v[5], v[6], v[7]);
    }
    else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            ((double2 *)p)[k] = make_double2(v[2 * k], v[2 * k + 1]);
    }
}

// PF = 1: the next block's stream is loaded into registers while the current
// block is reduced; PF = 0: no register stage (fewer VGPRs, more waves).
// The metadata of up to DICT_META blocks is loaded into LDS in one step, so
// the stream loads of a block wait on no global load of their own.
template <int PF, typename T, typename Epi>
__device__ inline void dict_spmv(const DictView &m, const T *__restrict__ va,
                                 const T *__restrict__ x, Epi epi)
{
    constexpr int NCH = DictStage<T>::NCH;
    static_assert(phg::DICT_BLOCK_ROWS <= BS, "one work-item per row");
    __shared__ __align__(16) T prod[phg::DICT_BLOCK_NNZ];
    __shared__ int32_t dsh[256];
    __shared__ SweepMeta md;
    int64_t first, end, step;
    phg::xcd_sweep(m.nblk, first, end, step);
    const int tid = threadIdx.x;
    for (int64_t c0 = first; c0 < end; c0 += step * DICT_META) {
        const int nb = sweep_meta(m, c0, end, step, md);
        DictStage<T> cur, nxt;
        sweep_fetch(m, md, va, x, 0, cur);
        for (int i = 0; i < nb; i++) {
            if (!PF && i > 0)
                sweep_fetch(m, md, va, x, i, cur);
            dsh[tid] = cur.d;
            lds_barrier();
            T xg[NCH][8];
#pragma unroll
            for (int k = 0; k < NCH; k++) {
                const int32_t e = cur.f0 + (tid + BS * k) * 8;
                if (e < cur.f1) {
                    const uint32_t w[4] = {cur.ib[k].x, cur.ib[k].y,
                                           cur.ib[k].z, cur.ib[k].w};
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const uint32_t u =
                            (w[j >> 1] >> (16 * (j & 1))) & 0xffffu;
                        xg[k][j] = x[dsh[u >> 8] + (int32_t)(u & 255u)];
                    }
                }
            }
            if (PF && i + 1 < nb)
                sweep_fetch(m, md, va, x, i + 1, nxt);
#pragma unroll
            for (int k = 0; k < NCH; k++) {
                const int32_t e = cur.f0 + (tid + BS * k) * 8;
                if (e < cur.f1) {
                    // group q of this wavefront's 512 entries, L entries per
                    // gather: entry j of lane l is j * L + l of the group
                    const int32_t q0 = ((e - cur.f0) / phg::DICT_GROUP) * phg::DICT_GROUP;
                    const int32_t L = min(phg::DICT_GROUP, cur.f1 - cur.f0 - q0) / 8;
                    const int l = (tid & 63);
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        prod[q0 + j * L + l] = cur.vb[k].v[j] * xg[k][j];
                }
            }
            lds_barrier();
            if (tid < cur.q1 - cur.q0) {
                T sum = ordered_sum(prod, cur.s0 - cur.f0, cur.s1 - cur.f0);
                epi(cur.q0 + tid, sum, cur.xr);
            }
            if (PF)
                cur = nxt;
        }
    }
}

// Staged-line CSR stream (the shift-4 dictionary layout): per block the
// distinct 16-column lines of x its entries use (at most XS_SLOTS) are read
// into LDS with coalesced loads; an entry's 16-bit code slot << 4 | offset
// is then its LDS index, so the per-entry gathers of x go to LDS instead of
// the vector L1.  Stream loads, product stage and ordered row sums as in
// dict_spmv.  PF = 1 loads the next block's stream into registers while the
// current block is reduced.
struct XsView {
    const int32_t *rp, *brow, *bent, *bdict, *dict;
    const uint16_t *idx;
    int64_t nblk, n;
};
template <int PF> struct XsViewT : XsView {};

// x line of a dictionary slot: the LW values from column d, 0 past n
template <typename T> struct XLine { T v[1 << phg::XS_SHIFT]; };
template <typename T>
__device__ inline void load_xline(const T *__restrict__ x, int64_t n, int32_t d,
                                  XLine<T> &xl)
{
    constexpr int LW = 1 << phg::XS_SHIFT;
    if ((int64_t)d + LW <= n) {
        if (sizeof(T) == 4) {
#pragma unroll
            for (int k = 0; k < LW / 4; k++) {
                const float4 a = ((const float4 *)(x + d))[k];
                xl.v[4 * k] = a.x; xl.v[4 * k + 1] = a.y;
                xl.v[4 * k + 2] = a.z; xl.v[4 * k + 3] = a.w;
            }
        }
        else {
#pragma unroll
            for (int k = 0; k < LW / 2; k++) {
                const double2 a = ((const double2 *)(x + d))[k];
                xl.v[2 * k] = a.x; xl.v[2 * k + 1] = a.y;
            }
        }
    }
    else {
#pragma unroll
        for (int k = 0; k < LW; k++)
            xl.v[k] = (int64_t)d + k < n ? x[d + k] : (T)0;
    }
}

template <int PF, typename T, typename Epi>
__device__ inline void xs_spmv(const XsView &m, const T *__restrict__ va,
                               const T *__restrict__ x, Epi epi)
{
    using Stage = DictStage<T, phg::XS_BLOCK_NNZ>;
    constexpr int NCH = Stage::NCH;
    constexpr int LW = 1 << phg::XS_SHIFT;
    static_assert(phg::XS_SLOTS <= BS, "one work-item per line");
    __shared__ __align__(16) T prod[phg::XS_BLOCK_NNZ];
    __shared__ __align__(16) T xsh[phg::XS_SLOTS * LW];
    __shared__ SweepMeta md;
    const DictView &dv = reinterpret_cast<const DictView &>(m);
    int64_t first, end, step;
    phg::xcd_sweep(m.nblk, first, end, step);
    const int tid = threadIdx.x;
    for (int64_t c0 = first; c0 < end; c0 += step * DICT_META) {
        const int nb = sweep_meta(dv, c0, end, step, md);
        Stage cur, nxt;
        // work-item t < nd holds the x line of dictionary slot t; with PF
        // the next block's lines load while this block's products run
        XLine<T> xl;
        sweep_fetch(dv, md, va, x, 0, cur);
        if (tid < md.d1[0] - md.d0[0])
            load_xline(x, m.n, cur.d, xl);
        for (int i = 0; i < nb; i++) {
            const int nd = md.d1[i] - md.d0[i];
            if (!PF && i > 0) {
                sweep_fetch(dv, md, va, x, i, cur);
                if (tid < nd)
                    load_xline(x, m.n, cur.d, xl);
            }
            // the previous block's products (the last xsh reads) ended at
            // its second barrier
            if (tid < nd) {
#pragma unroll
                for (int k = 0; k < LW; k++)
                    xsh[tid * LW + k] = xl.v[k];
            }
            if (PF && i + 1 < nb)
                sweep_fetch(dv, md, va, x, i + 1, nxt);
            lds_barrier();
#pragma unroll
            for (int k = 0; k < NCH; k++) {
                const int32_t e = cur.f0 + (tid + BS * k) * 8;
                if (e < cur.f1) {
                    const int32_t q0 = ((e - cur.f0) / phg::DICT_GROUP) * phg::DICT_GROUP;
                    const int32_t L = min(phg::DICT_GROUP, cur.f1 - cur.f0 - q0) / 8;
                    const int l = tid & 63;
                    const uint32_t w[4] = {cur.ib[k].x, cur.ib[k].y, cur.ib[k].z,
                                           cur.ib[k].w};
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const uint32_t u = (w[j >> 1] >> (16 * (j & 1))) & 0xffffu;
                        prod[q0 + j * L + l] = cur.vb[k].v[j] * xsh[u];
                    }
                }
            }
            if (PF && i + 1 < nb && tid < md.d1[i + 1] - md.d0[i + 1])
                load_xline(x, m.n, nxt.d, xl);
            lds_barrier();
            if (tid < cur.q1 - cur.q0) {
                T sum = ordered_sum(prod, cur.s0 - cur.f0, cur.s1 - cur.f0);
                epi(cur.q0 + tid, sum, cur.xr);
            }
            if (PF)
                cur = nxt;
        }
    }
}

template <typename T, int PF, typename Epi>
__device__ inline void spmv_apply(const XsViewT<PF> &m,
                                  const T *__restrict__ va,
                                  const T *__restrict__ x, Epi epi)
{
    xs_spmv<PF, T>(m, va, x, epi);
}

template <int PF> struct DictViewT : DictView {};
template <typename T, int PF, typename Epi>
__device__ inline void spmv_apply(const DictViewT<PF> &m,
                                  const T *__restrict__ va,
                                  const T *__restrict__ x, Epi epi)
{
    dict_spmv<PF, T>(m, va, x, epi);
}

// SELL construction: entries of permuted row i (original row perm[i]) with
// column indices renumbered through iperm; padding repeats the row's own
// index with value 0
__global__ void k_sell_fill(int64_t n, int64_t nslice, const int32_t *rp,
                            const int32_t *ci, const double *va,
                            const double *diag, const int32_t *perm,
                            const int32_t *iperm, const int64_t *sptr,
                            const int32_t *swidth, int32_t *scols,
                            double *svals, double *sdiag)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslice * 64)
        return;
    const int64_t sl = i / 64, lane = i % 64;
    const int64_t base = sptr[sl] + lane;
    const int w = swidth[sl];
    int k = 0;
    int32_t self = 0;
    if (i < n) {
        const int32_t r = perm[i];
        const int32_t j0 = rp[r], len = rp[r + 1] - j0;
        for (; k < len; k++) {
            scols[base + 64 * (int64_t)k] = iperm[ci[j0 + k]];
            svals[base + 64 * (int64_t)k] = va[j0 + k];
        }
        self = (int32_t)i;
        sdiag[i] = diag[r];
    }
    for (; k < w; k++) {
        scols[base + 64 * (int64_t)k] = self;
        svals[base + 64 * (int64_t)k] = 0.0;
    }
}

__global__ void k_gather(int64_t n, const int32_t *perm, const double *src,
                         double *dst)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        dst[i] = src[perm[i]];
}

__global__ void k_scatter(int64_t n, const int32_t *perm, const double *src,
                          double *dst)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        dst[perm[i]] = src[i];
}

// FP64 vector in PHG numbering <-> FP32 vector in the solver numbering
__global__ void k_gather_f32(int64_t n, const int32_t *perm, const double *src,
                             float *dst)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        dst[i] = (float)src[perm ? perm[i] : i];
}

__global__ void k_scatter_f32(int64_t n, const int32_t *perm, const float *src,
                              double *dst)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        dst[perm ? perm[i] : i] = (double)src[i];
}

//------------------------------------------------------------- plain SpMV
__global__ void __launch_bounds__(BS) k_spmv64(const int32_t *rp,
                                               const int32_t *ci,
                                               const double *va,
                                               const int32_t *rblk,
                                               int64_t nrblk, const double *x,
                                               double *y, double alpha,
                                               double beta)
{
    stream_spmv<double, NNZ_CAP, 0>(rp, ci, va, x, rblk, nrblk,
                                    [&](int32_t row, double s, double xr) {
                            // matvec.c:608-633 beta first, then alpha * A x
                            double yo = beta == 0.0 ? 0.0 : beta * y[row];
                            y[row] = yo + alpha * s;
                        });
}

// 4 consecutive elements with 16-byte accesses (float4, or two double2)
template <typename T> struct Q4 { T v[4]; };
template <typename T>
__device__ inline Q4<T> ld4(const T *p, int64_t q)
{
    Q4<T> r;
    if (sizeof(T) == 4) {
        const float4 a = ((const float4 *)p)[q];
        r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
    }
    else {
        const double2 a = ((const double2 *)p)[2 * q], b = ((const double2 *)p)[2 * q + 1];
        r.v[0] = a.x; r.v[1] = a.y; r.v[2] = b.x; r.v[3] = b.y;
    }
    return r;
}
template <typename T>
__device__ inline void st4(T *p, int64_t q, const Q4<T> &r)
{
    if (sizeof(T) == 4)
        ((float4 *)p)[q] = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
    else {
        ((double2 *)p)[2 * q] = make_double2(r.v[0], r.v[1]);
        ((double2 *)p)[2 * q + 1] = make_double2(r.v[2], r.v[3]);
    }
}

// calls f(i) for every element index of [0, n) this work-item owns, the
// groups of 4 through fq(q) (elements 4q .. 4q+3), the n % 4 tail elements
// through f(i) on the first work-group
template <typename FQ, typename F>
__device__ inline void for_each4(int64_t n, FQ fq, F f)
{
    const int64_t n4 = n >> 2;
    for (int64_t q = (int64_t)blockIdx.x * BS + threadIdx.x; q < n4;
         q += (int64_t)gridDim.x * BS)
        fq(q);
    if (blockIdx.x == 0 && (int64_t)threadIdx.x < n - 4 * n4)
        f(4 * n4 + threadIdx.x);
}

//------------------------------------------------------------------ BLAS-1
__global__ void __launch_bounds__(BS) k_dot(int64_t n, const double *x,
                                            const double *y, double *partials)
{
    double v[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        v[0] += x[i] * y[i];
    write_partials<1>(v, partials);
}

__global__ void __launch_bounds__(BS) k_final_sum(const double *partials,
                                                  int np, double *out)
{
    double o[1];
    sum_partials<1>(partials, np, o);
    if (threadIdx.x == 0)
        *out = o[0];
}

// solver-phg.c:207 jacobi_pc: x = b ./ diag, and the number of zero
// diagonal entries in the partials
__global__ void __launch_bounds__(BS) k_jacobi(int64_t n, const double *diag,
                                               const double *b, double *x,
                                               double *partials)
{
    double v[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        const double d = diag[i];
        if (d == 0.0)
            v[0] += 1.0;
        else
            x[i] = b[i] / d;
    }
    write_partials<1>(v, partials);
}

__global__ void __launch_bounds__(BS) k_axpby(int64_t n, double a,
                                              const double *x, double b,
                                              double *y)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        double t = b == 0.0 ? 0.0 : b * y[i];
        y[i] = a == 0.0 ? t : t + a * x[i];
    }
}

//--------------------------------------------------------- PCG kernels
// Prologue (solver-phg.c:505-533): |b| and |x|_1 partials ...
__global__ void __launch_bounds__(BS) k_pcg_init_norms(int64_t n,
                                                       const double *b,
                                                       const double *x,
                                                       double *part_out)
{
    double v[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        v[0] += b[i] * b[i];
        v[1] += fabs(x[i]);
    }
    write_partials<2>(v, part_out);
}

// ... then r = b - A x when |x|_1/n is not zero (else r = b; x = 0 when b
// is zero) and the partials (r,r), (r,z), #zero diagonals for iteration 0.
template <typename View>
__global__ void __launch_bounds__(BS) k_pcg_init_res(
    int64_t n, View mv, const double *va, const double *diag, const double *b,
    double *x, double *r, const double *part_in, int np_in, double *part_out,
    State *st0)
{
    double nb[2];
    sum_partials<2>(part_in, np_in, nb);
    const double rb = sqrt(nb[0]);
    const bool bzero = is_zero(rb);
    const bool xnz = !is_zero(nb[1] / (double)n);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st0->rb = rb;
        st0->xnz = xnz;
    }
    double v[3] = {0.0, 0.0, 0.0};
    auto acc = [&](int64_t row, double ri) {
        r[row] = ri;
        double d = diag[row];
        v[0] += ri * ri;
        if (d == 0.0)
            v[2] += 1.0;
        else
            v[1] += ri * (ri / d);
    };
    if (xnz && !bzero) {
        spmv_apply<double>(mv, va, x,
                           [&](int32_t row, double s, double xr) { acc(row, -s + b[row]); });
    }
    else {
        for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * BS) {
            if (bzero)
                x[i] = 0.0;
            acc(i, bzero ? 0.0 : b[i]);
        }
    }
    write_partials<3>(v, part_out);
}

// p-update with the loop control of solver-phg.c:535-559
template <typename T>
__global__ void __launch_bounds__(BS) k_pcg_p(int64_t n, const T *r,
                                              const T *diag, T *p,
                                              const double *part_in, int np_in,
                                              const State *sin, State *sout)
{
    const State s = *sin;
    const bool w0 = blockIdx.x == 0 && threadIdx.x == 0;
    if (s.done) {
        if (w0)
            *sout = s;
        return;
    }
    // the work-item's first group of 4 rows is loaded while the partials
    // are summed
    const int64_t n4 = n >> 2, q0 = (int64_t)blockIdx.x * BS + threadIdx.x;
    const bool has0 = q0 < n4;
    Q4<T> r0, d0, p0;
    double t[3];                             // (r,r), (r,z), #zero diagonals
    sum_partials_then<3>(part_in, np_in, t, [&]() {
        if (has0) {
            r0 = ld4(r, q0);
            d0 = ld4(diag, q0);
            p0 = ld4((const T *)p, q0);
        }
    });
    const double res = sqrt(t[0]), rz = t[1];
    State o = s;
    o.res = res;
    bool stop = false;
    double beta = 0.0;
    if (s.nits == 0) {
        if (is_zero(s.rb)) {
            o.res = 0.0;
            o.done = CONVERGED;
            stop = true;
        }
        else {
            double tol = res * s.rtol;       // get_tol (solver-phg.c:104)
            if (tol < s.rb * s.btol)
                tol = s.rb * s.btol;
            if (tol < s.atol)
                tol = s.atol;
            o.tol = tol;
            o.r0 = res;
            if (!(res > tol && s.maxit > 0)) {
                o.done = CONVERGED;
                stop = true;
            }
            else if (t[2] != 0.0) {          // solver-phg.c:233
                o.done = ZERO_DIAG;
                stop = true;
            }
            else {
                o.nits = 1;
                o.rz = rz;
                o.rz_prev = 0.0;
            }
        }
    }
    else {
        if (!(res > s.tol && s.nits < s.maxit)) {
            o.done = CONVERGED;
            stop = true;
        }
        else if (s.rz == 0.0) {              // solver-phg.c:547
            o.nits = s.nits + 1;
            o.done = BREAKDOWN;
            stop = true;
        }
        else {
            o.nits = s.nits + 1;
            o.rz_prev = s.rz;
            o.rz = rz;
            beta = rz / s.rz;
        }
    }
    if (w0)
        *sout = o;
    if (stop)
        return;
    const bool first = s.nits == 0;
    const T tb = (T)beta;
    auto quad = [&](int64_t q, const Q4<T> &rv, const Q4<T> &dv, Q4<T> pv) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const T z = rv.v[k] / dv.v[k];
            pv.v[k] = first ? z : tb * pv.v[k] + z;
        }
        st4(p, q, pv);
    };
    if (has0)
        quad(q0, r0, d0, p0);
    for (int64_t q = q0 + (int64_t)gridDim.x * BS; q < n4; q += (int64_t)gridDim.x * BS)
        quad(q, ld4(r, q), ld4(diag, q), first ? Q4<T>{{T(0), T(0), T(0), T(0)}}
                                               : ld4((const T *)p, q));
    if (blockIdx.x == 0 && (int64_t)threadIdx.x < n - 4 * n4) {
        const int64_t i = 4 * n4 + threadIdx.x;
        T z = r[i] / diag[i];
        p[i] = first ? z : tb * p[i] + z;
    }
}

template <typename T, typename View>
__global__ void __launch_bounds__(BS) k_pcg_spmv(
    View mv, const T *va, const T *p, T *Ap, double *part_out, const State *st)
{
    if (st->done)
        return;
    double v[1] = {0.0};
    spmv_apply<T>(mv, va, p, [&](int32_t row, T s, T xr) {
        Ap[row] = s;
        v[0] += (double)xr * (double)s;     // xr = p[row]
    });
    write_partials<1>(v, part_out);
}

template <typename T>
__global__ void __launch_bounds__(BS) k_pcg_update(int64_t n, T *x,
                                                   const T *p, T *r,
                                                   const T *Ap, const T *diag,
                                                   const double *part_in,
                                                   int np_in, double *part_out,
                                                   State *st)
{
    if (st->done)
        return;
    // the work-item's first group of 4 rows is loaded while the partials
    // of (p, Ap) are summed
    const int64_t n4 = n >> 2, q0 = (int64_t)blockIdx.x * BS + threadIdx.x;
    const bool has0 = q0 < n4;
    Q4<T> x0, p0, r0, a0, d0;
    double pap[1];
    sum_partials_then<1>(part_in, np_in, pap, [&]() {
        if (has0) {
            x0 = ld4((const T *)x, q0);
            p0 = ld4(p, q0);
            r0 = ld4((const T *)r, q0);
            a0 = ld4(Ap, q0);
            d0 = ld4(diag, q0);
        }
    });
    if (pap[0] == 0.0) {                     // solver-phg.c:561
        if (blockIdx.x == 0 && threadIdx.x == 0)
            st->done = BREAKDOWN;
        return;
    }
    const T alpha = (T)(st->rz / pap[0]);
    double v[3] = {0.0, 0.0, 0.0};
    auto quad = [&](int64_t q, Q4<T> xv, const Q4<T> &pv, Q4<T> rv, const Q4<T> &av,
                    const Q4<T> &dv) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            xv.v[k] += alpha * pv.v[k];
            const T ri = rv.v[k] - alpha * av.v[k];
            rv.v[k] = ri;
            const T z = ri / dv.v[k];
            v[0] += (double)ri * (double)ri;
            v[1] += (double)ri * (double)z;
        }
        st4(x, q, xv);
        st4(r, q, rv);
    };
    if (has0)
        quad(q0, x0, p0, r0, a0, d0);
    for (int64_t q = q0 + (int64_t)gridDim.x * BS; q < n4; q += (int64_t)gridDim.x * BS)
        quad(q, ld4((const T *)x, q), ld4(p, q), ld4((const T *)r, q), ld4(Ap, q),
             ld4(diag, q));
    if (blockIdx.x == 0 && (int64_t)threadIdx.x < n - 4 * n4) {
        const int64_t i = 4 * n4 + threadIdx.x;
        x[i] += alpha * p[i];
        T ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        T z = ri / diag[i];
        v[0] += (double)ri * (double)ri;
        v[1] += (double)ri * (double)z;
    }
    write_partials<3>(v, part_out);
}

//----------------------------------------------- mixed-precision refinement
// r = b - A x in FP64; partials of |r|^2 and |b|^2
template <typename View>
__global__ void __launch_bounds__(BS) k_ir_residual(
    View mv, const double *va, const double *b, const double *x, double *r,
    double *part_out)
{
    double v[2] = {0.0, 0.0};
    spmv_apply<double>(mv, va, x, [&](int32_t row, double s, double xr) {
        double bi = b[row], ri = bi - s;
        r[row] = ri;
        v[0] += ri * ri;
        v[1] += bi * bi;
    });
    write_partials<2>(v, part_out);
}

// one work-group: totals into out[0..1] = (|r|, |b|)
__global__ void __launch_bounds__(BS) k_ir_norms(const double *partials,
                                                 int np, double *out)
{
    double t[2];
    sum_partials<2>(partials, np, t);
    if (threadIdx.x == 0) {
        out[0] = sqrt(t[0]);
        out[1] = sqrt(t[1]);
    }
}

// inner right-hand side r32 = r * scale, x32 = 0; partials as k_pcg_init_res
__global__ void __launch_bounds__(BS) k_ir_inner_init(
    int64_t n, const double *r, double scale, const float *diag32, float *r32,
    float *x32, double *part_out)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        float ri = (float)(r[i] * scale);
        r32[i] = ri;
        x32[i] = 0.0f;
        float d = diag32[i];
        v[0] += (double)ri * (double)ri;
        if (d == 0.0f)
            v[2] += 1.0;
        else
            v[1] += (double)ri * (double)(ri / d);
    }
    write_partials<3>(v, part_out);
}

// x += s * d  (FP64 correction)
__global__ void __launch_bounds__(BS) k_ir_correct(int64_t n, double *x,
                                                   const float *d, double s)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        x[i] += s * (double)d[i];
}

//------------------------------------------------------------------ BiCGstab
// solver-phg.c:607 solver_BiCGstab with jacobi_pc, FP64, one grid pass per
// step of the reference's loop:
//   k_bi_p      loop control (residual test, rho1 == 0), beta,
//               p = r + beta (p - omega v), ph = p / diag
//   k_bi_spmv1  v = A ph, partials of (rh, v)
//   k_bi_s      alpha = rho1 / (rh, v), s = r - alpha v, sh = s / diag,
//               partials of (s, s)
//   k_bi_spmv2  |s| <= 1e-60 exit (flag), else t = A sh, partials of
//               (t, s), (t, t)
//   k_bi_update omega = (t, s) / (t, t), ph = alpha ph + omega sh,
//               x += ph, r = s - omega t, partials of (r, r), (r, rh);
//               on the exit flag x += alpha ph (the caller then forms the
//               true residual, solver-phg.c:718-727)
// Every kernel reads the whole state at its start and work-group 0 writes
// the next copy (two copies alternate per launch), so no work-group reads
// a value another one writes in the same launch.
struct BiState {
    double rb, tol, res, rho0, rho1, alpha, omega, beta;
    double rtol, btol, atol;
    int32_t nits, done, maxit, small;
};
constexpr int BI_SMALL_S = 4;    // done: the |s| <= 1e-60 exit was taken

__global__ void __launch_bounds__(BS) k_bi_p(int64_t n, const double *r,
                                             const double *diag, double *p,
                                             double *ph, const double *v,
                                             const double *part_in, int np_init,
                                             int np_iter, const BiState *sin,
                                             BiState *sout)
{
    const BiState s = *sin;
    const bool w0 = blockIdx.x == 0 && threadIdx.x == 0;
    if (s.done) {
        if (w0)
            *sout = s;
        return;
    }
    double t[3];                 // (r,r), (r,rh), #zero diagonals
    sum_partials<3>(part_in, s.nits == 0 ? np_init : np_iter, t);
    const double res = sqrt(t[0]);
    BiState o = s;
    o.res = res;
    bool stop = false;
    if (s.nits == 0) {
        if (is_zero(s.rb)) {
            o.res = 1.0;         // solver-phg.c:631: the caller's value
            o.done = CONVERGED;
            stop = true;
        }
        else {
            double tol = res * s.rtol;       // get_tol (solver-phg.c:104)
            if (tol < s.rb * s.btol)
                tol = s.rb * s.btol;
            if (tol < s.atol)
                tol = s.atol;
            o.tol = tol;
            if (!(res > tol && s.maxit > 0)) {
                o.done = CONVERGED;
                stop = true;
            }
            else if (t[2] != 0.0) {          // solver-phg.c:233
                o.done = ZERO_DIAG;
                stop = true;
            }
        }
    }
    else if (!(res > s.tol && s.nits < s.maxit)) {
        o.done = CONVERGED;
        stop = true;
    }
    double beta = 0.0;
    if (!stop) {
        o.nits = s.nits + 1;
        const double rho1 = t[1];
        if (rho1 == 0.0) {                   // solver-phg.c:685
            o.done = BREAKDOWN;
            stop = true;
        }
        else {
            if (o.nits > 1)
                beta = (rho1 * s.alpha) / (s.rho0 * s.omega);
            o.beta = beta;
            o.rho1 = rho1;
            o.rho0 = rho1;
        }
    }
    if (w0)
        *sout = o;
    if (stop)
        return;
    const bool first = o.nits == 1;
    const double mo = -s.omega;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        double pi;
        if (first)
            pi = r[i];
        else {
            const double si = p[i] + mo * v[i];
            pi = r[i] + beta * si;
        }
        p[i] = pi;
        ph[i] = pi / diag[i];
    }
}

template <typename View>
__global__ void __launch_bounds__(BS) k_bi_spmv1(View mv, const double *va,
                                                 const double *ph, double *v,
                                                 const double *rh, double *part_out,
                                                 const BiState *sin, BiState *sout)
{
    const BiState s = *sin;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *sout = s;
    if (s.done)
        return;
    double a[1] = {0.0};
    spmv_apply<double>(mv, va, ph, [&](int32_t row, double sum, double) {
        v[row] = sum;
        a[0] += rh[row] * sum;
    });
    write_partials<1>(a, part_out);
}

__global__ void __launch_bounds__(BS) k_bi_s(int64_t n, const double *r,
                                             const double *v, const double *diag,
                                             double *sg, double *sh,
                                             const double *part_in, int np_in,
                                             double *part_out, const BiState *sin,
                                             BiState *sout)
{
    const BiState s = *sin;
    const bool w0 = blockIdx.x == 0 && threadIdx.x == 0;
    if (s.done) {
        if (w0)
            *sout = s;
        return;
    }
    double rv[1];
    sum_partials<1>(part_in, np_in, rv);
    BiState o = s;
    o.alpha = s.rho1 / rv[0];
    if (w0)
        *sout = o;
    const double ma = -o.alpha;
    double a[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        const double si = r[i] + ma * v[i];
        sg[i] = si;
        sh[i] = si / diag[i];
        a[0] += si * si;
    }
    write_partials<1>(a, part_out);
}

template <typename View>
__global__ void __launch_bounds__(BS) k_bi_spmv2(View mv, const double *va,
                                                 const double *sh, double *tg,
                                                 const double *sg,
                                                 const double *part_in, int np_in,
                                                 double *part_out,
                                                 const BiState *sin, BiState *sout)
{
    const BiState s = *sin;
    const bool w0 = blockIdx.x == 0 && threadIdx.x == 0;
    if (s.done) {
        if (w0)
            *sout = s;
        return;
    }
    double ss[1];
    sum_partials<1>(part_in, np_in, ss);
    BiState o = s;
    o.small = sqrt(ss[0]) <= 1e-60 ? 1 : 0;  // solver-phg.c:717
    if (w0)
        *sout = o;
    if (o.small)
        return;
    double a[2] = {0.0, 0.0};
    spmv_apply<double>(mv, va, sh, [&](int32_t row, double sum, double) {
        tg[row] = sum;
        a[0] += sum * sg[row];
        a[1] += sum * sum;
    });
    write_partials<2>(a, part_out);
}

__global__ void __launch_bounds__(BS) k_bi_update(int64_t n, double *x, double *ph,
                                                  const double *sh, double *r,
                                                  const double *sg, const double *tg,
                                                  const double *rh,
                                                  const double *part_in, int np_in,
                                                  double *part_out,
                                                  const BiState *sin, BiState *sout)
{
    const BiState s = *sin;
    const bool w0 = blockIdx.x == 0 && threadIdx.x == 0;
    if (s.done) {
        if (w0)
            *sout = s;
        return;
    }
    if (s.small) {
        BiState o = s;
        o.done = BI_SMALL_S;
        if (w0)
            *sout = o;
        for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * BS)
            x[i] += s.alpha * ph[i];
        return;
    }
    double t[2];
    sum_partials<2>(part_in, np_in, t);
    BiState o = s;
    o.omega = t[0] / t[1];
    if (w0)
        *sout = o;
    const double al = s.alpha, om = o.omega, mo = -o.omega;
    double a[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        const double pi = al * ph[i] + om * sh[i];
        ph[i] = pi;
        x[i] += pi;
        const double ri = sg[i] + mo * tg[i];
        r[i] = ri;
        a[0] += ri * ri;
        a[1] += ri * rh[i];
    }
    write_partials<3>(a, part_out);
}

// prologue: r = b - A x (or b when x is zero; x = 0 when b is zero),
// rh = r, partials of (r,r), (r,rh) = (r,r) and the zero diagonals
template <typename View>
__global__ void __launch_bounds__(BS) k_bi_init(int64_t n, View mv, const double *va,
                                                const double *diag, const double *b,
                                                double *x, double *r, double *rh,
                                                const double *part_in, int np_in,
                                                double *part_out, BiState *st0)
{
    double nb[2];
    sum_partials<2>(part_in, np_in, nb);
    const double rb = sqrt(nb[0]);
    const bool bzero = is_zero(rb);
    const bool xnz = !is_zero(nb[1] / (double)n);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        st0->rb = rb;
    double a[3] = {0.0, 0.0, 0.0};
    auto acc = [&](int64_t row, double ri) {
        r[row] = ri;
        rh[row] = ri;
        a[0] += ri * ri;
        a[1] += ri * ri;
        if (diag[row] == 0.0)
            a[2] += 1.0;
    };
    if (xnz && !bzero)
        spmv_apply<double>(mv, va, x,
                           [&](int32_t row, double sum, double) { acc(row, -sum + b[row]); });
    else
        for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * BS) {
            if (bzero)
                x[i] = 0.0;
            acc(i, bzero ? 0.0 : b[i]);
        }
    write_partials<3>(a, part_out);
}

// true residual of the |s| <= 1e-60 exit: partials of (b - A x)^2
template <typename View>
__global__ void __launch_bounds__(BS) k_bi_true_res(View mv, const double *va,
                                                    const double *b, const double *x,
                                                    double *part_out)
{
    double a[1] = {0.0};
    spmv_apply<double>(mv, va, x, [&](int32_t row, double sum, double) {
        const double ri = -sum + b[row];
        a[0] += ri * ri;
    });
    write_partials<1>(a, part_out);
}

//------------------------------------------------------------------- GMRES
// The GMRES family of solver-phg.c:59-77 (gmres_l, gmres_r, fgmres_r,
// lgmres_l, lgmres_r, flgmres_r) with jacobi_pc: the vector work of one
// Arnoldi step runs on the device --
//   k_gm_pc        flv_i = src / diag (right variants: src = v_i or an
//                  augmentation vector)
//   k_gm_spmv      v_{i+1} = A flv_i (right) or (A src) / diag (left),
//                  partials of (v_{i+1}, v_1)
//   k_gm_mgs       modified Gram-Schmidt step k: h_k from the partials,
//                  v_{i+1} -= h_k v_k, partials of (v_{i+1}, v_{k+1}) or of
//                  |v_{i+1}|^2 after the last k
//   k_gm_scale     h_{i+1} = |v_{i+1}|, v_{i+1} /= h_{i+1} (not on a zero
//                  h_{i+1}), and the next step's flv = v_{i+1} / diag
// -- and the host reads the column h_1..h_{i+1} and applies the Givens
// rotations, least-squares solve and restart logic of the reference in
// FP64.  Work-group 0 writes each h_k to hcol.
constexpr int GM_MAXV = 4096;        // largest subspace (restart + augk) + 1

__global__ void __launch_bounds__(BS) k_zero(int64_t n, double *x)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        x[i] = 0.0;
}

__global__ void __launch_bounds__(BS) k_gm_pc(int64_t n, const double *src,
                                              const double *diag, double *dst)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        dst[i] = src[i] / diag[i];
}

template <typename View>
__global__ void __launch_bounds__(BS) k_gm_spmv(View mv, const double *va,
                                                const double *f, double *v,
                                                const double *v1, const double *ldiag,
                                                double *part_out)
{
    double a[1] = {0.0};
    spmv_apply<double>(mv, va, f, [&](int32_t row, double s, double) {
        const double w = ldiag ? s / ldiag[row] : s;
        v[row] = w;
        a[0] += w * v1[row];
    });
    write_partials<1>(a, part_out);
}

__global__ void __launch_bounds__(BS) k_gm_mgs(int64_t n, double *w, const double *vk,
                                               const double *vnext,
                                               const double *part_in, int np_in,
                                               double *part_out, double *hcol, int k)
{
    double h[1];
    sum_partials<1>(part_in, np_in, h);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        hcol[k] = h[0];
    const double mh = -h[0];
    double a[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        const double wi = w[i] + mh * vk[i];
        w[i] = wi;
        a[0] += wi * (vnext ? vnext[i] : wi);
    }
    write_partials<1>(a, part_out);
}

__global__ void __launch_bounds__(BS) k_gm_scale(int64_t n, double *w,
                                                 const double *part_in, int np_in,
                                                 double *hcol, int k,
                                                 const double *diag, double *fnext)
{
    double h2[1];
    sum_partials<1>(part_in, np_in, h2);
    const double h = sqrt(h2[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        hcol[k] = h;
    if (is_zero(h))
        return;                      // happy breakdown: the host stops
    const double sc = 1.0 / h;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        const double wi = sc * w[i];
        w[i] = wi;
        if (fnext)
            fnext[i] = wi / diag[i];
    }
}

// out = sum_j c_j src_j in the reference's order (c_1 src_1, then + c_j
// src_j) over two runs of vectors stored n apart: src_j = b1 + j n for
// j < m1, then b2 + (j - m1) n; c_j = coef[j], or cval for every j when coef
// is null.  Then z = out (z non-null), x += out / pdiag (pdiag non-null) or
// x += out (x non-null), zp = out (zp non-null).  Each element's sources
// are read before any output is written, so an output may be a source.
__global__ void __launch_bounds__(BS) k_gm_combo(int64_t n, const double *b1, int m1,
                                                 const double *b2, int m2,
                                                 const double *coef, double cval,
                                                 double *z, double *x,
                                                 const double *pdiag, double *zp)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        double o = (coef ? coef[0] : cval) * b1[i];
        for (int j = 1; j < m1; j++)
            o += (coef ? coef[j] : cval) * b1[(size_t)j * n + i];
        for (int j = 0; j < m2; j++)
            o += (coef ? coef[m1 + j] : cval) * b2[(size_t)j * n + i];
        if (z)
            z[i] = o;
        if (x)
            x[i] = x[i] + (pdiag ? o / pdiag[i] : o);
        if (zp)
            zp[i] = o;
    }
}

// z = b - A x (or b when x is zero), partials of |z|^2 and of the number of
// zero diagonals; with left set, z = (b - A x) / diag instead and a third
// partial of |z|^2 (the first stays |b - A x|^2)
template <typename View>
__global__ void __launch_bounds__(BS) k_gm_resid(int64_t n, View mv, const double *va,
                                                 const double *diag, const double *b,
                                                 const double *x, int xnz, int left,
                                                 double *z, double *part_out)
{
    double a[3] = {0.0, 0.0, 0.0};
    auto acc = [&](int64_t row, double ri) {
        a[0] += ri * ri;
        const double di = diag[row];
        if (di == 0.0)
            a[1] += 1.0;
        if (left) {
            const double zi = ri / di;
            z[row] = zi;
            a[2] += zi * zi;
        }
        else
            z[row] = ri;
    };
    if (xnz)
        spmv_apply<double>(mv, va, x,
                           [&](int32_t row, double s, double) { acc(row, b[row] + (-s)); });
    else
        for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * BS)
            acc(i, b[i]);
    write_partials<3>(a, part_out);
}

//------------------------------------------ reliable-update mixed precision
// FP32 PCG whose residual is replaced by the FP64 residual b - A x (x kept
// in FP64) whenever the FP32 residual has dropped by the factor delta below
// its largest value since the previous replacement, or has reached the
// tolerance.  The search direction is kept across replacements, so the
// Krylov recurrence is not restarted; convergence is only declared on an
// FP64 residual.  One slot of launches per step:
//   k_ru_p       decides the slot kind; ITER: p = z + beta p;
//                UPDATE: x += x32, x32 = 0
//   k_ru_spmv    ITER only: Ap = A32 p, partials of (p, Ap)
//   k_ru_update  ITER: x32 += alpha p, r32 -= alpha Ap, partials;
//                UPDATE: r32 = b - A x in FP64 with the solver CSR, partials
// so an ITER slot runs no kernel that only exits.  A layout with its own
// row order (SELL) adds k_ru_resid, which computes the UPDATE residual
// through the layout instead.  k_ru_p of the next slot reads the partials
// of the kernel that wrote them: after ITER k_ru_update (grid g), after
// UPDATE k_ru_update or k_ru_resid (grid g or gs), after the prologue
// k_ru_init (grid gs).

// prologue: r32 = b - A x (or b when x is zero), x32 = 0, partials of
// (r,r), (r,z), #zero diagonals computed from the FP64 residual
template <typename View>
__global__ void __launch_bounds__(BS) k_ru_init(
    int64_t n, View mv, const double *va, const double *diag, const double *b,
    double *x, float *r32, float *x32, const double *part_in, int np_in,
    double *part_out, State *st0)
{
    double nb[2];
    sum_partials<2>(part_in, np_in, nb);
    const double rb = sqrt(nb[0]);
    const bool bzero = is_zero(rb);
    const bool xnz = !is_zero(nb[1] / (double)n);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st0->rb = rb;
        st0->xnz = xnz;
    }
    double v[3] = {0.0, 0.0, 0.0};
    auto acc = [&](int64_t row, double ri) {
        r32[row] = (float)ri;
        x32[row] = 0.0f;
        double d = diag[row];
        v[0] += ri * ri;
        if (d == 0.0)
            v[2] += 1.0;
        else
            v[1] += ri * (ri / d);
    };
    if (xnz && !bzero) {
        spmv_apply<double>(mv, va, x,
                           [&](int32_t row, double s, double xr) { acc(row, -s + b[row]); });
    }
    else {
        for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * BS) {
            if (bzero)
                x[i] = 0.0;
            acc(i, bzero ? 0.0 : b[i]);
        }
    }
    write_partials<3>(v, part_out);
}

__global__ void __launch_bounds__(BS) k_ru_p(int64_t n, const float *r32,
                                             const float *diag32, float *p32,
                                             double *x, float *x32,
                                             const double *part_in, int np_iter,
                                             int np_upd, int np_start,
                                             const State *sin, State *sout)
{
    const State s = *sin;
    const bool w0 = blockIdx.x == 0 && threadIdx.x == 0;
    if (s.done) {
        if (w0)
            *sout = s;
        return;
    }
    // this work-item's first group of 4 rows of an iteration slot is loaded
    // while the partials are summed (an update slot leaves them unused)
    const int64_t n4 = n >> 2, q0 = (int64_t)blockIdx.x * BS + threadIdx.x;
    const bool has0 = q0 < n4;
    Q4<float> r0, d0, p0;
    double t[3];
    sum_partials_then<3>(part_in,
                         s.mode == RU_ITER ? np_iter : (s.mode == RU_UPDATE ? np_upd : np_start),
                         t, [&]() {
                             if (has0) {
                                 r0 = ld4(r32, q0);
                                 d0 = ld4(diag32, q0);
                                 p0 = ld4((const float *)p32, q0);
                             }
                         });
    const double res = sqrt(t[0]), rz = t[1];
    State o = s;
    o.res = res;
    int kind = RU_ITER;           // what this slot does
    double beta = 0.0;
    bool first = false;
    if (s.mode == RU_START) {
        // r0 = b - A x in FP64: tolerance as get_tol (solver-phg.c:104)
        if (is_zero(s.rb)) {
            o.res = 0.0;
            o.done = CONVERGED;
        }
        else {
            double tol = res * s.rtol;
            if (tol < s.rb * s.btol)
                tol = s.rb * s.btol;
            if (tol < s.atol)
                tol = s.atol;
            o.tol = tol;
            o.r0 = res;
            o.rmax = res;
            if (!(res > tol && s.maxit > 0))
                o.done = CONVERGED;
            else if (t[2] != 0.0)
                o.done = ZERO_DIAG;
            else {
                o.nits = 1;
                o.rz = rz;
                first = true;
            }
        }
    }
    else if (s.mode == RU_UPDATE) {
        // the residual was just replaced by the FP64 one
        o.rmax = res;
        if (!(res > s.tol))
            o.done = CONVERGED;
        else if (s.nits >= s.maxit)
            o.done = CONVERGED;
        else if (s.rz == 0.0)
            o.done = BREAKDOWN;
        else {
            o.nits = s.nits + 1;
            o.rz = rz;
            beta = rz / s.rz;
        }
    }
    else {
        if (!(res > s.tol) || res < s.delta * s.rmax || s.nits >= s.maxit) {
            kind = RU_UPDATE;       // fold x32 into x, then FP64 residual
            o.updates = s.updates + 1;
        }
        else if (s.rz == 0.0)
            o.done = BREAKDOWN;
        else {
            o.nits = s.nits + 1;
            o.rz = rz;
            o.rmax = res > s.rmax ? res : s.rmax;
            beta = rz / s.rz;
        }
    }
    o.mode = kind;
    if (w0)
        *sout = o;
    if (o.done)
        return;
    if (kind == RU_UPDATE) {
        for_each4(n, [&](int64_t q) {
            Q4<double> xv = ld4((const double *)x, q);
            Q4<float> dv = ld4((const float *)x32, q);
#pragma unroll
            for (int k = 0; k < 4; k++)
                xv.v[k] += (double)dv.v[k];
            st4(x, q, xv);
            st4(x32, q, Q4<float>{{0.0f, 0.0f, 0.0f, 0.0f}});
        }, [&](int64_t i) {
            x[i] += (double)x32[i];
            x32[i] = 0.0f;
        });
        return;
    }
    const float tb = (float)beta;
    auto quad = [&](int64_t q, const Q4<float> &rv, const Q4<float> &dv, Q4<float> pv) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float z = rv.v[k] / dv.v[k];
            pv.v[k] = first ? z : tb * pv.v[k] + z;
        }
        st4(p32, q, pv);
    };
    if (has0)
        quad(q0, r0, d0, p0);
    for (int64_t q = q0 + (int64_t)gridDim.x * BS; q < n4; q += (int64_t)gridDim.x * BS)
        quad(q, ld4(r32, q), ld4(diag32, q),
             first ? Q4<float>{{0.0f, 0.0f, 0.0f, 0.0f}} : ld4((const float *)p32, q));
    if (blockIdx.x == 0 && (int64_t)threadIdx.x < n - 4 * n4) {
        const int64_t i = 4 * n4 + threadIdx.x;
        float z = r32[i] / diag32[i];
        p32[i] = first ? z : tb * p32[i] + z;
    }
}

template <typename View>
__global__ void __launch_bounds__(BS) k_ru_spmv(View mv, const float *va,
                                                const float *p, float *Ap,
                                                double *part_out,
                                                const State *st)
{
    if (st->done || st->mode != RU_ITER)
        return;
    double v[1] = {0.0};
    spmv_apply<float>(mv, va, p, [&](int32_t row, float s, float xr) {
        Ap[row] = s;
        v[0] += (double)xr * (double)s;     // xr = p[row]
    });
    write_partials<1>(v, part_out);
}

// FP64 residual of a replacement slot: r32 = b - A x with the solver CSR
// (row blocks of the streaming SpMV, rows summed in storage order),
// partials of (r,r), (r,z) and the number of zero diagonals as k_ru_init
// writes them
__device__ inline void ru_residual64(const int32_t *__restrict__ rp,
                                     const int32_t *__restrict__ ci,
                                     const double *__restrict__ va,
                                     const int32_t *__restrict__ rblk,
                                     int64_t nrblk,
                                     const double *__restrict__ diag,
                                     const double *__restrict__ b,
                                     const double *__restrict__ x, float *r32,
                                     double *part_out)
{
    double v[3] = {0.0, 0.0, 0.0};
    stream_spmv<double, phg::SPMV_BLOCK_NNZ, 0>(
        rp, ci, va, x, rblk, nrblk, [&](int32_t row, double s, double) {
            const double ri = b[row] - s;
            r32[row] = (float)ri;
            const double d = diag[row];
            v[0] += ri * ri;
            if (d == 0.0)
                v[2] += 1.0;
            else
                v[1] += ri * (ri / d);
        });
    write_partials<3>(v, part_out);
}

// ITER: x32 += alpha p, r32 -= alpha Ap, partials; UPDATE: the FP64
// residual (x was folded by k_ru_p)
__global__ void __launch_bounds__(BS) k_ru_update(int64_t n, float *x32,
                                                  const float *p, float *r,
                                                  const float *Ap,
                                                  const float *diag,
                                                  const double *part_in,
                                                  int np_in, double *part_out,
                                                  State *st, const int32_t *rp,
                                                  const int32_t *ci,
                                                  const double *va,
                                                  const int32_t *rblk,
                                                  int64_t nrblk,
                                                  const double *diag64,
                                                  const double *b,
                                                  const double *x,
                                                  int resid_here,
                                                  const uint32_t *fxc,
                                                  const float *fpart)
{
    if (st->done)
        return;
    if (st->mode == RU_UPDATE) {
        if (resid_here)
            ru_residual64(rp, ci, va, rblk, nrblk, diag64, b, x, r, part_out);
        return;
    }
    if (st->mode != RU_ITER)
        return;
    // this work-item's first group of 4 rows is loaded while the partials
    // of (p, Ap) are summed (none of it depends on alpha)
    const int64_t n4 = n >> 2, q0 = (int64_t)blockIdx.x * BS + threadIdx.x;
    const bool has0 = q0 < n4;
    Q4<float> x0, p0, r0, a0, d0;
    uint4 c0 = make_uint4(0u, 0u, 0u, 0u);
    double pap[1];
    sum_partials_then<1>(part_in, np_in, pap, [&]() {
        if (has0) {
            x0 = ld4((const float *)x32, q0);
            p0 = ld4(p, q0);
            r0 = ld4((const float *)r, q0);
            a0 = ld4(Ap, q0);
            d0 = ld4(diag, q0);
            if (fxc)
                c0 = ((const uint4 *)fxc)[q0];
        }
    });
    if (pap[0] == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            st->done = BREAKDOWN;
        return;
    }
    const float alpha = (float)(st->rz / pap[0]);
    double v[3] = {0.0, 0.0, 0.0};
    // Ap of a row shared between element blocks of the matrix-free
    // operator: the sum of its partials in block order (code: partial start
    // | count << 24)
    auto shared = [&](uint32_t c, float a) {
        if (c == phg::EOP_NOT_SHARED)
            return a;
        const uint32_t q0 = c & 0xffffffu, q1 = q0 + (c >> 24);
        float sum = 0.0f;
        for (uint32_t q = q0; q < q1; q++)
            sum += fpart[q];
        return sum;
    };
    auto one = [&](int64_t i) {
        x32[i] += alpha * p[i];
        const float api = fxc ? shared(fxc[i], Ap[i]) : Ap[i];
        float ri = r[i] - alpha * api;
        r[i] = ri;
        float z = ri / diag[i];
        v[0] += (double)ri * (double)ri;
        v[1] += (double)ri * (double)z;
    };
    // for_each4's order: the work-item's groups of 4, then the tail
    auto quad = [&](int64_t q, Q4<float> xv, const Q4<float> &pv, Q4<float> rv, Q4<float> av,
                    const Q4<float> &dv, const uint4 &cv) {
        if (fxc) {
            av.v[0] = shared(cv.x, av.v[0]);
            av.v[1] = shared(cv.y, av.v[1]);
            av.v[2] = shared(cv.z, av.v[2]);
            av.v[3] = shared(cv.w, av.v[3]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            xv.v[k] += alpha * pv.v[k];
            const float ri = rv.v[k] - alpha * av.v[k];
            rv.v[k] = ri;
            const float z = ri / dv.v[k];
            v[0] += (double)ri * (double)ri;
            v[1] += (double)ri * (double)z;
        }
        st4(x32, q, xv);
        st4(r, q, rv);
    };
    if (has0)
        quad(q0, x0, p0, r0, a0, d0, c0);
    for (int64_t q = q0 + (int64_t)gridDim.x * BS; q < n4; q += (int64_t)gridDim.x * BS)
        quad(q, ld4((const float *)x32, q), ld4(p, q), ld4((const float *)r, q), ld4(Ap, q),
             ld4(diag, q), fxc ? ((const uint4 *)fxc)[q] : make_uint4(0u, 0u, 0u, 0u));
    if (blockIdx.x == 0 && (int64_t)threadIdx.x < n - 4 * n4)
        one(4 * n4 + threadIdx.x);
    write_partials<3>(v, part_out);
}

// UPDATE slots of layouts with their own row order (SELL): r32 = b - A x
// in FP64 through the layout, partials
template <typename View>
__global__ void __launch_bounds__(BS) k_ru_resid(View mv, const double *va,
                                                 const double *diag,
                                                 const double *b,
                                                 const double *x, float *r32,
                                                 double *part_out,
                                                 const State *st)
{
    if (st->done || st->mode != RU_UPDATE)
        return;
    double v[3] = {0.0, 0.0, 0.0};
    spmv_apply<double>(mv, va, x, [&](int32_t row, double s, double xr) {
        double ri = b[row] - s;
        r32[row] = (float)ri;
        double d = diag[row];
        v[0] += ri * ri;
        if (d == 0.0)
            v[2] += 1.0;
        else
            v[1] += ri * (ri / d);
    });
    write_partials<3>(v, part_out);
}

// x += x32 (final fold of the FP32 increment)
__global__ void __launch_bounds__(BS) k_ru_fold(int64_t n, double *x,
                                                const float *x32)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        x[i] += (double)x32[i];
}

// plain copy, the reference point for bandwidth and HBM-counter scaling
__global__ void __launch_bounds__(BS) k_copy(int64_t n, const double *a,
                                             double *b)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        b[i] = a[i];
}

// *st = h (the solver state's start values, passed by value)
__global__ void k_state_set(State *st, State h)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *st = h;
}

__global__ void k_to_f32(int64_t n, const double *a, float *b)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        b[i] = (float)a[i];
}

//------------------------------------------------------------- host side
// partial-sum segments (each MAX_GRID * 4 doubles)
enum { SEG_NORMS = 0, SEG_RES, SEG_S, SEG_U, SEG_IR, SEG_N };
// storage formats of the solver SpMV
enum {
    FMT_B2K = 0,      // CsrBlockView<2048, 0>
    FMT_B4K,          // CsrBlockView<4096, 0>
    FMT_B2K_PAR,      // CsrBlockView<2048, 2>
    FMT_B4K_PAR,      // CsrBlockView<4096, 2>
    FMT_WAVE,         // CsrWaveView
    FMT_SELL,         // SellView
    FMT_DICT,         // DictViewT<1>
    FMT_DICT_NP,      // DictViewT<0>
    FMT_XS,           // XsViewT<0>
    FMT_XS_PF,        // XsViewT<1>
    FMT_N
};
static_assert(FMT_N == PHG_HIP_FMT_COUNT, "format list and C ABI disagree");
using CsrB2K = CsrBlockView<2048, 0>;
using CsrB4K = CsrBlockView<4096, 0>;
using CsrB2KPar = CsrBlockView<2048, 2>;
using CsrB4KPar = CsrBlockView<4096, 2>;

// calls f with a value of the view type of format fmt
template <typename Fn>
int with_format(int fmt, Fn &&f)
{
    switch (fmt) {
    case FMT_B2K: return f(CsrB2K{});
    case FMT_B4K: return f(CsrB4K{});
    case FMT_B2K_PAR: return f(CsrB2KPar{});
    case FMT_B4K_PAR: return f(CsrB4KPar{});
    case FMT_WAVE: return f(CsrWaveView{});
    case FMT_DICT: return f(DictViewT<1>{});
    case FMT_DICT_NP: return f(DictViewT<0>{});
    case FMT_XS: return f(XsViewT<0>{});
    case FMT_XS_PF: return f(XsViewT<1>{});
    default: return f(SellView{});
    }
}

struct Workspace {
    int64_t n = 0;
    double *r = nullptr, *p = nullptr, *Ap = nullptr, *partials = nullptr;
    double *xp = nullptr, *bp = nullptr;  // x and b in the SELL numbering
    double *scal = nullptr;               // small device scalars
    float *r32 = nullptr, *p32 = nullptr, *Ap32 = nullptr, *x32 = nullptr;
    State *st = nullptr;                  // [2] ping-pong copies + 1 spare
    // BiCGstab: rh, ph, v, s, sh, t (one allocation) and its states
    double *bi = nullptr;
    BiState *bst = nullptr;
    // GMRES basis, preconditioned and augmentation vectors (gm_cnt of n);
    // combination coefficients (2 slots of gm_ccap, staged through pinned
    // host memory) and a pinned slot for reading a column of H back
    double *gm = nullptr;
    int gm_cnt = 0;
    double *gm_coef = nullptr, *gm_coef_h = nullptr;
    int gm_ccap = 0;
    double *seg(int k) { return partials + (size_t)k * 4 * MAX_GRID; }
};

int cap_grid()
{
    int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    return std::min(cus * 8, MAX_GRID);
}

int grid_for(int64_t n)
{
    int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    int64_t g = (n + BS - 1) / BS;
    int64_t cap = std::min<int64_t>((int64_t)cus * 4, MAX_GRID);
    return (int)std::max<int64_t>(1, std::min(g, cap));
}

// views and launch grids of the three formats
template <int CAP, int MODE>
CsrBlockView<CAP, MODE> view_of(const phg_hip_mat *A, CsrBlockView<CAP, MODE> *)
{
    if (CAP == phg::SPMV_LARGE_NNZ)
        return CsrBlockView<CAP, MODE>{A->q_rowptr, A->q_cols, A->q_rblk1k,
                                       A->q_nrblk1k};
    return CsrBlockView<CAP, MODE>{A->q_rowptr, A->q_cols, A->q_rblk,
                                   A->q_nrblk};
}
CsrWaveView view_of(const phg_hip_mat *A, CsrWaveView *)
{
    return CsrWaveView{A->q_rowptr, A->q_cols, A->q_wblk, A->q_nwblk};
}
template <int PF>
DictViewT<PF> view_of(const phg_hip_mat *A, DictViewT<PF> *)
{
    DictViewT<PF> v;
    v.rp = A->cd.rowptr;
    v.brow = A->cd.blk_row;
    v.bent = A->cd.blk_ent;
    v.bdict = A->cd.blk_dict;
    v.dict = A->cd.dict;
    v.idx = A->cd.idx;
    v.nblk = A->cd.nblk;
    return v;
}
template <int PF>
int64_t units_of(const phg_hip_mat *A, DictViewT<PF> *) { return A->cd.nblk; }
template <int PF>
XsViewT<PF> view_of(const phg_hip_mat *A, XsViewT<PF> *)
{
    XsViewT<PF> v;
    v.rp = A->xs.rowptr;
    v.brow = A->xs.blk_row;
    v.bent = A->xs.blk_ent;
    v.bdict = A->xs.blk_dict;
    v.dict = A->xs.dict;
    v.idx = A->xs.idx;
    v.nblk = A->xs.nblk;
    v.n = A->n;
    return v;
}
template <int PF>
int64_t units_of(const phg_hip_mat *A, XsViewT<PF> *) { return A->xs.nblk; }
SellView view_of(const phg_hip_mat *A, SellView *)
{
    return SellView{A->d_sptr, A->d_swidth, A->d_scols, A->nslice, A->n};
}
PhgPermView view_of(const phg_hip_mat *A, PhgPermView *)
{
    return PhgPermView{A->d_rowptr, A->d_cols, A->q_perm, A->q_iperm, A->n};
}
int64_t units_of(const phg_hip_mat *A, PhgPermView *)
{
    return (A->n + PPV_ROWS - 1) / PPV_ROWS;
}
template <int CAP, int MODE>
int64_t units_of(const phg_hip_mat *A, CsrBlockView<CAP, MODE> *)
{
    return CAP == phg::SPMV_LARGE_NNZ ? A->q_nrblk1k : A->q_nrblk;
}
int64_t units_of(const phg_hip_mat *A, CsrWaveView *)
{
    return (A->q_nwblk + BS / 64 - 1) / (BS / 64) * (BS / 64);
}
int64_t units_of(const phg_hip_mat *A, SellView *)
{
    return (A->nslice + BS / 64 - 1) / (BS / 64);
}

// For the wavefront stream the work-group range is taken over wavefront
// blocks (waves interleave inside it), so the grid is sized by groups of 4.
// The grid of an SpMV kernel is at most the number of work-groups the chip
// holds at once for that kernel (its occupancy times the CU count), so every
// work-group starts in the first wave of dispatch and walks its contiguous
// range of blocks; a larger grid would run a second, partly empty round.
int resident_groups(const void *kern)
{
    static std::vector<std::pair<const void *, int>> cache;
    for (auto &c : cache)
        if (c.first == kern)
            return c.second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BS, 0) !=
            hipSuccess ||
        per_cu <= 0) {
        (void)hipGetLastError();   // the query's failure is not the caller's
        per_cu = 1;
    }
    int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    int g = std::min(per_cu * cus, MAX_GRID);
    cache.emplace_back(kern, g);
    return g;
}

template <typename View, typename K>
int grid_spmv(const phg_hip_mat *A, K kern)
{
    int64_t u = units_of(A, (View *)nullptr);
    if (std::is_same<View, CsrWaveView>::value)
        u /= BS / 64;
    const int cap = resident_groups((const void *)kern);
    return (int)std::max<int64_t>(1, std::min<int64_t>(u, cap));
}

template <typename View> struct FmtData {
    static const double *v64(const phg_hip_mat *A) { return A->q_vals; }
    static const float *v32(const phg_hip_mat *A) { return A->q_vals32; }
    static const double *d64(const phg_hip_mat *A) { return A->q_diag; }
    static const float *d32(const phg_hip_mat *A) { return A->q_diag32; }
    static const int32_t *perm(const phg_hip_mat *A) { return A->q_perm; }
};
template <int PF> struct FmtData<XsViewT<PF>> {
    static const double *v64(const phg_hip_mat *A) { return A->xs.val64; }
    static const float *v32(const phg_hip_mat *A) { return A->xs.val32; }
    static const double *d64(const phg_hip_mat *A) { return A->q_diag; }
    static const float *d32(const phg_hip_mat *A) { return A->q_diag32; }
    static const int32_t *perm(const phg_hip_mat *A) { return A->q_perm; }
};
template <int PF> struct FmtData<DictViewT<PF>> {
    static const double *v64(const phg_hip_mat *A) { return A->cd.val64; }
    static const float *v32(const phg_hip_mat *A) { return A->cd.val32; }
    static const double *d64(const phg_hip_mat *A) { return A->q_diag; }
    static const float *d32(const phg_hip_mat *A) { return A->q_diag32; }
    static const int32_t *perm(const phg_hip_mat *A) { return A->q_perm; }
};
template <> struct FmtData<PhgPermView> {
    static const double *v64(const phg_hip_mat *A) { return A->d_vals; }
    static const float *v32(const phg_hip_mat *) { return nullptr; }
    static const double *d64(const phg_hip_mat *A) { return A->q_diag; }
    static const float *d32(const phg_hip_mat *A) { return A->q_diag32; }
    static const int32_t *perm(const phg_hip_mat *A) { return A->q_perm; }
};
template <> struct FmtData<SellView> {
    static const double *v64(const phg_hip_mat *A) { return A->d_svals; }
    static const float *v32(const phg_hip_mat *A) { return A->d_svals32; }
    static const double *d64(const phg_hip_mat *A) { return A->d_sdiag; }
    static const float *d32(const phg_hip_mat *A) { return A->d_sdiag32; }
    static const int32_t *perm(const phg_hip_mat *A) { return A->d_perm; }
};

void release_ws(Workspace *w)
{
    phg::dfree(w->r); phg::dfree(w->p); phg::dfree(w->Ap);
    phg::dfree(w->xp); phg::dfree(w->bp);
    phg::dfree(w->partials); phg::dfree(w->st); phg::dfree(w->scal);
    phg::dfree(w->r32); phg::dfree(w->p32); phg::dfree(w->Ap32);
    phg::dfree(w->x32);
    phg::dfree(w->bi);
    phg::dfree(w->bst);
    phg::dfree(w->gm);
    phg::dfree(w->gm_coef);
    if (w->gm_coef_h)
        (void)hipHostFree(w->gm_coef_h);
    *w = Workspace();
}

// Solver vectors of a matrix: each matrix owns its workspace, so solves of
// different matrices can be in flight at the same time (on different
// streams); solves of one matrix are ordered.
Workspace *get_ws(const phg_hip_mat *A_, int need32)
{
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    Workspace *w = (Workspace *)A->ws;
    if (w == nullptr) {
        w = new Workspace();
        A->ws = w;
    }
    if (w->n != A->n) {
        release_ws(w);
        size_t n8 = (size_t)A->n * sizeof(double);
        if (phg::dmalloc((void **)&w->r, n8) || phg::dmalloc((void **)&w->p, n8) ||
            phg::dmalloc((void **)&w->Ap, n8) || phg::dmalloc((void **)&w->xp, n8) ||
            phg::dmalloc((void **)&w->bp, n8) ||
            phg::dmalloc((void **)&w->partials,
                         (size_t)SEG_N * 4 * MAX_GRID * sizeof(double)) ||
            phg::dmalloc((void **)&w->scal, 16 * sizeof(double)) ||
            phg::dmalloc((void **)&w->st, 3 * sizeof(State))) {
            release_ws(w);
            return nullptr;
        }
        w->n = A->n;
    }
    if (need32 && w->r32 == nullptr) {
        size_t n4 = (size_t)A->n * sizeof(float);
        if (phg::dmalloc((void **)&w->r32, n4) || phg::dmalloc((void **)&w->p32, n4) ||
            phg::dmalloc((void **)&w->Ap32, n4) || phg::dmalloc((void **)&w->x32, n4))
            return nullptr;
    }
    return w;
}

// Optional per-kernel timing with HIP events recorded on the library stream
// around each launch of the PCG kernels (phg_hip_profile).
enum { KP_P64 = 0, KP_SPMV64, KP_UPD64, KP_P32, KP_SPMV32, KP_UPD32, KP_N };
struct KernelProfile {
    bool on = false;
    std::vector<hipEvent_t> ev[KP_N];   // start/stop pairs
    size_t used[KP_N] = {};
    int64_t launches[KP_N] = {};
    double ms[KP_N] = {};
};
KernelProfile g_prof;

inline void prof_mark(int kind, hipStream_t s)
{
    std::vector<hipEvent_t> &v = g_prof.ev[kind];
    size_t &u = g_prof.used[kind];
    if (u >= v.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess)
            return;
        v.push_back(e);
    }
    (void)hipEventRecord(v[u++], s);
}

// folds the recorded pairs into the totals (the stream must be idle)
void prof_collect()
{
    for (int k = 0; k < KP_N; k++) {
        std::vector<hipEvent_t> &v = g_prof.ev[k];
        for (size_t i = 0; i + 1 < g_prof.used[k]; i += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, v[i], v[i + 1]) == hipSuccess) {
                g_prof.ms[k] += ms;
                g_prof.launches[k] += 1;
            }
        }
        g_prof.used[k] = 0;
    }
}

// one PCG iteration (3 kernels) with launch index k on the given precision;
// part_in holds the residual partials of the previous iteration (or of the
// prologue, np_first entries, when k == 0)
template <typename T, typename View>
void launch_iteration(const phg_hip_mat *A, const T *va, const T *diag,
                      T *x, T *r, T *p, T *Ap, Workspace *ws, State *st,
                      int64_t k, int np_first, hipStream_t s)
{
    int64_t n = A->n;
    int g = grid_for(n), gs = grid_spmv<View>(A, k_pcg_spmv<T, View>);
    const int np_in = k == 0 ? np_first : g;
    State *sin = st + (k & 1), *sout = st + ((k + 1) & 1);
    double *pres = ws->seg(SEG_RES), *ps = ws->seg(SEG_S);
    const int base = sizeof(T) == 8 ? KP_P64 : KP_P32;
    const bool prof = g_prof.on;
    if (prof) prof_mark(base + 0, s);
    hipLaunchKernelGGL(k_pcg_p<T>, dim3(g), dim3(BS), 0, s, n, (const T *)r,
                       diag, p, (const double *)pres, np_in,
                       (const State *)sin, sout);
    if (prof) { prof_mark(base + 0, s); prof_mark(base + 1, s); }
    hipLaunchKernelGGL((k_pcg_spmv<T, View>), dim3(gs), dim3(BS), 0, s,
                       view_of(A, (View *)nullptr), va, (const T *)p, Ap, ps,
                       (const State *)sout);
    if (prof) { prof_mark(base + 1, s); prof_mark(base + 2, s); }
    hipLaunchKernelGGL(k_pcg_update<T>, dim3(g), dim3(BS), 0, s, n, x,
                       (const T *)p, r, (const T *)Ap, diag,
                       (const double *)ps, gs, pres, sout);
    if (prof) prof_mark(base + 2, s);
}

int read_state(State *d, State *h, hipStream_t s)
{
    PHG_HIP_CHECK(hipMemcpyAsync(h, d, sizeof(State), hipMemcpyDeviceToHost, s));
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    return PHG_HIP_OK;
}

int check_solve_args(const phg_hip_mat *A, const phg_hip_vec *b,
                     const phg_hip_vec *x)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr || b == nullptr || x == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null argument");
    if (b->n != A->n || x->n != A->n)
        return phg::set_error(PHG_HIP_E_ARG, "size mismatch (%lld, %lld, %lld)",
                              (long long)A->n, (long long)b->n,
                              (long long)x->n);
    if (!A->eliminated || A->d_diag == nullptr)
        return phg::set_error(PHG_HIP_E_ARG,
                              "matrix not assembled: call phg_hip_mat_eliminate");
    return PHG_HIP_OK;
}

// Runs PCG iterations (launch indices from 0) until the state copy reports
// completion; *out receives the final state.
template <typename T, typename View>
int run_pcg_loop(const phg_hip_mat *A, const T *va, const T *diag, T *x, T *r,
                 T *p, T *Ap, Workspace *ws, int maxit, int np_first,
                 int first_chunk, State *out)
{
    hipStream_t s = phg::ctx()->stream;
    State *hm = (State *)phg::ctx()->host_mailbox;
    int64_t k = 0;
    int chunk = g_prof.on ? 1 : first_chunk;
    int rc;
    for (;;) {
        for (int i = 0; i < chunk; i++, k++)
            launch_iteration<T, View>(A, va, diag, x, r, p, Ap, ws, ws->st, k,
                                      np_first, s);
        PHG_HIP_CHECK(hipGetLastError());
        if ((rc = read_state(ws->st + (k & 1), hm, s)))
            return rc;
        if (hm->done != RUNNING)
            break;
        // the device stops at maxit; a state that never finishes means the
        // kernels did not run
        if (k > (int64_t)maxit + 256)
            return phg::set_error(PHG_HIP_E_HIP, "PCG state did not advance");
        if (!g_prof.on)
            chunk = std::min(chunk * 2, 64);
    }
    if (g_prof.on)
        prof_collect();
    *out = *hm;
    return PHG_HIP_OK;
}

int ensure_format(phg_hip_mat *A, int fmt, int need32);

// solver-phg.c:607 solver_BiCGstab in the numbering of the format
template <typename View>
int bicgstab_impl(phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x,
                  int32_t maxit, double rtol, double btol, double atol,
                  double *residual)
{
    using F = FmtData<View>;
    hipStream_t s = phg::ctx()->stream;
    Workspace *ws = get_ws(A, 0);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "solver workspace allocation failed");
    const int64_t n = A->n;
    if (ws->bi == nullptr &&
        (phg::dmalloc((void **)&ws->bi, (size_t)6 * n * sizeof(double)) ||
         phg::dmalloc((void **)&ws->bst, 2 * sizeof(BiState))))
        return phg::set_error(PHG_HIP_E_NOMEM, "BiCGstab workspace allocation failed");
    double *rh = ws->bi, *ph = rh + n, *v = ph + n, *sg = v + n, *sh = sg + n,
           *tg = sh + n;
    BiState h;
    memset(&h, 0, sizeof(h));
    h.rtol = rtol;
    h.btol = btol;
    h.atol = atol;
    h.maxit = maxit;
    PHG_HIP_CHECK(hipMemcpyAsync(ws->bst, &h, sizeof(BiState), hipMemcpyHostToDevice, s));
    PHG_HIP_CHECK(hipStreamSynchronize(s));     // h lives on this stack
    const int g = grid_for(n), gs = grid_spmv<View>(A, k_bi_spmv1<View>);
    const View mv = view_of(A, (View *)nullptr);
    const double *va = F::v64(A), *d = F::d64(A);
    const double *bb = b->d;
    double *xx = x->d;
    const int32_t *perm = F::perm(A);
    if (perm) {
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)b->d, ws->bp);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)x->d, ws->xp);
        bb = ws->bp;
        xx = ws->xp;
    }
    double *pn = ws->seg(SEG_NORMS), *pr = ws->seg(SEG_RES), *pv = ws->seg(SEG_S),
           *pss = ws->seg(SEG_U), *pt = ws->seg(SEG_IR);
    hipLaunchKernelGGL(k_pcg_init_norms, dim3(g), dim3(BS), 0, s, n, bb,
                       (const double *)xx, pn);
    hipLaunchKernelGGL((k_bi_init<View>), dim3(gs), dim3(BS), 0, s, n, mv, va, d, bb,
                       xx, ws->r, rh, (const double *)pn, g, pr, ws->bst);
    PHG_HIP_CHECK(hipGetLastError());
    BiState *st = ws->bst;
    BiState *hm = (BiState *)phg::ctx()->host_mailbox;
    int64_t j = 0;                       // kernel launches so far
    auto sio = [&](int64_t k) { return std::make_pair(st + (k & 1), st + ((k + 1) & 1)); };
    int chunk = 8;
    for (int64_t it = 0;; ) {
        for (int c = 0; c < chunk; c++, it++) {
            auto p0 = sio(j++);
            hipLaunchKernelGGL(k_bi_p, dim3(g), dim3(BS), 0, s, n,
                               (const double *)ws->r, d, ws->p, ph, (const double *)v,
                               (const double *)pr, gs, g, (const BiState *)p0.first,
                               p0.second);
            auto p1 = sio(j++);
            hipLaunchKernelGGL((k_bi_spmv1<View>), dim3(gs), dim3(BS), 0, s, mv, va,
                               (const double *)ph, v, (const double *)rh, pv,
                               (const BiState *)p1.first, p1.second);
            auto p2 = sio(j++);
            hipLaunchKernelGGL(k_bi_s, dim3(g), dim3(BS), 0, s, n, (const double *)ws->r,
                               (const double *)v, d, sg, sh, (const double *)pv, gs, pss,
                               (const BiState *)p2.first, p2.second);
            auto p3 = sio(j++);
            hipLaunchKernelGGL((k_bi_spmv2<View>), dim3(gs), dim3(BS), 0, s, mv, va,
                               (const double *)sh, tg, (const double *)sg,
                               (const double *)pss, g, pt, (const BiState *)p3.first,
                               p3.second);
            auto p4 = sio(j++);
            hipLaunchKernelGGL(k_bi_update, dim3(g), dim3(BS), 0, s, n, xx, ph,
                               (const double *)sh, ws->r, (const double *)sg,
                               (const double *)tg, (const double *)rh,
                               (const double *)pt, gs, pr, (const BiState *)p4.first,
                               p4.second);
        }
        PHG_HIP_CHECK(hipGetLastError());
        PHG_HIP_CHECK(hipMemcpyAsync(hm, st + (j & 1), sizeof(BiState),
                                     hipMemcpyDeviceToHost, s));
        PHG_HIP_CHECK(hipStreamSynchronize(s));
        if (hm->done != RUNNING)
            break;
        if (it > (int64_t)maxit + 256)
            return phg::set_error(PHG_HIP_E_HIP, "BiCGstab state did not advance");
        chunk = std::min(chunk * 2, 64);
    }
    BiState fin = *hm;
    if (fin.done == BI_SMALL_S) {
        // solver-phg.c:720-726: residual of the updated x
        hipLaunchKernelGGL((k_bi_true_res<View>), dim3(gs), dim3(BS), 0, s, mv, va, bb,
                           (const double *)xx, pr);
        PHG_HIP_CHECK(hipGetLastError());
        std::vector<double> part((size_t)gs);
        PHG_HIP_CHECK(hipMemcpyAsync(part.data(), pr, (size_t)gs * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
        PHG_HIP_CHECK(hipStreamSynchronize(s));
        double sum = 0.0;
        for (double q : part)
            sum += q;
        fin.res = sqrt(sum);
    }
    if (perm)
        hipLaunchKernelGGL(k_scatter, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)ws->xp, x->d);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    if (residual)
        *residual = fin.res;
    if (fin.done == ZERO_DIAG)
        return phg::set_error(PHG_HIP_E_ZERO_DIAG, "zero diagonal in Jacobi preconditioner");
    return fin.nits;
}

// value k of np work-group partials, summed on the host in block order
int host_partial_sum(const double *d, int np, int k, hipStream_t s, double *out)
{
    std::vector<double> h((size_t)np);
    PHG_HIP_CHECK(hipMemcpyAsync(h.data(), d + (size_t)k * np, (size_t)np * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    double t = 0.0;
    for (int i = 0; i < np; i++)
        t += h[(size_t)i];
    *out = t;
    return PHG_HIP_OK;
}

// IsZero (utils.h:451) on the host
inline bool is_zero_host(double a) { return fabs(a) < 10.0 * 2.2250738585072014e-308; }

// The GMRES family of solver-phg.c:59-77 in the numbering of the format:
// type 0 gmres_l (solver-phg.c:783), 1 gmres_r (:1051), 2 fgmres_r (:1263),
// 3 lgmres_l (:1481), 4 lgmres_r (:1801), 5 flgmres_r (:2063).
template <typename View>
int gmres_impl(phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x, int type,
               int restart, int augk, int32_t maxit, double rtol, double btol,
               double atol, double *residual)
{
    using F = FmtData<View>;
    hipStream_t s = phg::ctx()->stream;
    if (type < PHG_HIP_GMRES_L || type > PHG_HIP_FLGMRES_R)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_gmres: unknown GMRES type %d", type);
    const bool left = type == PHG_HIP_GMRES_L || type == PHG_HIP_LGMRES_L;
    const bool flex = type == PHG_HIP_FGMRES_R || type == PHG_HIP_FLGMRES_R;
    const bool aug = type >= PHG_HIP_LGMRES_L;
    if (restart <= 0)
        restart = maxit + 1;
    if (!aug)
        augk = 0;                    // Unused(augk) in the plain variants
    else if (augk <= 0)
        augk = 2;
    const int mk = restart, m = restart + augk;
    if (m + 1 > GM_MAXV || augk > 64)
        return phg::set_error(PHG_HIP_E_ARG,
                              "phg_hip_gmres: restart + augk must be below %d and augk "
                              "at most 64", GM_MAXV);
    Workspace *ws = get_ws(A, 0);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "solver workspace allocation failed");
    const int64_t n = A->n;
    // v_1..v_{m+1}, flv_1..flv_m, zp_0..zp_{augk-1}, z
    const int cnt = (m + 1) + m + augk + 1;
    if (ws->gm_cnt < cnt) {
        phg::dfree(ws->gm);
        ws->gm = nullptr;
        ws->gm_cnt = 0;
        if (phg::dmalloc((void **)&ws->gm, (size_t)cnt * n * sizeof(double)))
            return phg::set_error(PHG_HIP_E_NOMEM, "GMRES workspace allocation failed");
        ws->gm_cnt = cnt;
    }
    if (ws->gm_ccap < m + 2) {
        phg::dfree(ws->gm_coef);
        if (ws->gm_coef_h)
            (void)hipHostFree(ws->gm_coef_h);
        ws->gm_coef = ws->gm_coef_h = nullptr;
        ws->gm_ccap = 0;
        if (phg::dmalloc((void **)&ws->gm_coef, 2 * (size_t)(m + 2) * sizeof(double)) ||
            hipHostMalloc((void **)&ws->gm_coef_h, 3 * (size_t)(m + 2) * sizeof(double),
                          hipHostMallocDefault) != hipSuccess) {
            ws->gm_coef_h = nullptr;
            return phg::set_error(PHG_HIP_E_NOMEM, "GMRES workspace allocation failed");
        }
        ws->gm_ccap = m + 2;
    }
    // coefficients of a combination into slot k; a slot is rewritten only
    // after a later stream synchronisation (every cycle ends with one)
    auto put_coef = [&](int k, const double *c, int cnt_c) -> const double * {
        double *h = ws->gm_coef_h + (size_t)k * ws->gm_ccap;
        double *dv = ws->gm_coef + (size_t)k * ws->gm_ccap;
        std::copy(c, c + cnt_c, h);
        if (hipMemcpyAsync(dv, h, (size_t)cnt_c * sizeof(double), hipMemcpyHostToDevice, s) !=
            hipSuccess)
            return nullptr;
        return dv;
    };
    auto V = [&](int j) { return ws->gm + (size_t)(j - 1) * n; };              // 1..m+1
    auto FL = [&](int j) { return ws->gm + (size_t)(m + j) * n; };             // 1..m
    auto ZP = [&](int j) { return ws->gm + (size_t)(2 * m + 1 + j) * n; };     // 0..augk-1
    double *z = ws->gm + (size_t)(2 * m + 1 + augk) * n;
    const int g = grid_for(n), gs = grid_spmv<View>(A, k_gm_spmv<View>);
    const View mv = view_of(A, (View *)nullptr);
    const double *va = F::v64(A), *d = F::d64(A);
    const double *bb = b->d;
    double *xx = x->d;
    const int32_t *perm = F::perm(A);
    if (perm) {
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)b->d, ws->bp);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)x->d, ws->xp);
        bb = ws->bp;
        xx = ws->xp;
    }
    double *pn = ws->seg(SEG_NORMS), *pr = ws->seg(SEG_RES);
    double *pp[2] = {ws->seg(SEG_S), ws->seg(SEG_U)};
    double *hcol = ws->scal;         // 16 doubles: h_1..h_{i+1} of a step
    if (m + 2 > 16) {
        // longer columns live in the partials segment after the MGS ones
        hcol = ws->seg(SEG_IR);
    }
    static_assert(GM_MAXV + 1 <= 4 * MAX_GRID, "a column of H fits a partials segment");
    int rc;
    double nb[2];
    hipLaunchKernelGGL(k_pcg_init_norms, dim3(g), dim3(BS), 0, s, n, bb,
                       (const double *)xx, pn);
    PHG_HIP_CHECK(hipGetLastError());
    if ((rc = host_partial_sum(pn, g, 0, s, &nb[0])) ||
        (rc = host_partial_sum(pn, g, 1, s, &nb[1])))
        return rc;
    const double b_norm = sqrt(nb[0]);
    std::vector<double> cbuf((size_t)(m + 2));
    std::vector<double> H((size_t)(m + 2) * (m + 2)), yv((size_t)(m + 2)),
        sv((size_t)(m + 2)), c1((size_t)(m + 2)), c2((size_t)(m + 2));
    auto Hm = [&](int a, int c) -> double & { return H[(size_t)a * (m + 2) + c]; };
    int nits = 0;
    double res = 1.0;                // GMRES_Solve's initial value
    // r = b - A x (x != 0, solver-phg.c:2158: phgVecNorm1(x) != 0); the
    // left variants keep z = M^-1 r, the right ones z = r
    auto residual_pass = [&](int xnz, double *rr, double *zz2, double *nzero) {
        hipLaunchKernelGGL((k_gm_resid<View>), dim3(gs), dim3(BS), 0, s, n, mv, va, d, bb,
                           (const double *)xx, xnz, left ? 1 : 0, z, pr);
        PHG_HIP_CHECK(hipGetLastError());
        double r2;
        int e;
        if ((e = host_partial_sum(pr, gs, 0, s, &r2)) ||
            (nzero && (e = host_partial_sum(pr, gs, 1, s, nzero))) ||
            (e = host_partial_sum(pr, gs, left ? 2 : 0, s, zz2)))
            return e;
        *rr = sqrt(r2);
        return PHG_HIP_OK;
    };
    if (is_zero_host(b_norm)) {
        hipLaunchKernelGGL(k_zero, dim3(g), dim3(BS), 0, s, n, xx);
    }
    else {
        double zz, nzero, r_norm;
        if ((rc = residual_pass(nb[1] != 0.0 ? 1 : 0, &r_norm, &zz, &nzero)))
            return rc;
        res = r_norm;
        double tol = r_norm * rtol;
        if (tol < b_norm * btol)
            tol = b_norm * btol;
        if (tol < atol)
            tol = atol;
        // jacobi_pc stops on a zero diagonal (solver-phg.c:230) the first
        // time it runs: at once for the left variants (M z = r0), in the
        // first cycle for the plain right ones (x += M^-1 V y), at the
        // first Arnoldi step for the flexible ones
        if (nzero != 0.0 &&
            (left || (maxit > 0 && (!flex || r_norm > tol))))
            return phg::set_error(PHG_HIP_E_ZERO_DIAG,
                                  "zero diagonal in Jacobi preconditioner");
        sv[0] = 1.0;
        double s0 = 1.0;
        bool false_conv = false, break_flag = false;
        int itr_out = 0;
        while (nits < maxit) {
            const int rs = itr_out <= augk ? mk + itr_out : mk + augk;
            sv[1] = r_norm = sqrt(zz);
            // v_1 = z / |z| (vaxpby(1 / r_norm, z, 0, v_1)); right: flv_1 = v_1 / diag
            hipLaunchKernelGGL(k_gm_combo, dim3(g), dim3(BS), 0, s, n, (const double *)z, 1,
                               (const double *)nullptr, 0, (const double *)nullptr,
                               1.0 / r_norm, V(1), (double *)nullptr,
                               (const double *)nullptr, (double *)nullptr);
            if (!left)
                hipLaunchKernelGGL(k_gm_pc, dim3(g), dim3(BS), 0, s, n,
                                   (const double *)V(1), d, FL(1));
            int i = 0;
            std::fill(H.begin(), H.end(), 0.0);
            while (i < rs && (r_norm > tol || false_conv) && nits < maxit) {
                i++, nits++;
                const double *src = i <= mk ? V(i) : ZP(i - mk - 1);
                if (!left && i > mk)
                    hipLaunchKernelGGL(k_gm_pc, dim3(g), dim3(BS), 0, s, n, src, d, FL(i));
                double *w = V(i + 1);
                hipLaunchKernelGGL((k_gm_spmv<View>), dim3(gs), dim3(BS), 0, s, mv, va,
                                   left ? src : (const double *)FL(i), w,
                                   (const double *)V(1), left ? d : (const double *)nullptr,
                                   pp[0]);
                for (int k = 1; k <= i; k++)
                    hipLaunchKernelGGL(k_gm_mgs, dim3(g), dim3(BS), 0, s, n, w,
                                       (const double *)V(k),
                                       k < i ? (const double *)V(k + 1) : (const double *)nullptr,
                                       (const double *)pp[(k - 1) & 1], k == 1 ? gs : g,
                                       pp[k & 1], hcol, k);
                // the next step's input is v_{i+1} unless it uses an
                // augmentation vector
                hipLaunchKernelGGL(k_gm_scale, dim3(g), dim3(BS), 0, s, n, w,
                                   (const double *)pp[i & 1], g, hcol, i + 1, d,
                                   !left && i + 1 <= mk ? FL(i + 1) : (double *)nullptr);
                PHG_HIP_CHECK(hipGetLastError());
                double *hm = ws->gm_coef_h + 2 * (size_t)ws->gm_ccap;
                PHG_HIP_CHECK(hipMemcpyAsync(hm, hcol, (size_t)(i + 2) * sizeof(double),
                                             hipMemcpyDeviceToHost, s));
                PHG_HIP_CHECK(hipStreamSynchronize(s));
                for (int k = 1; k <= i + 1; k++)
                    Hm(i, k) = hm[k];
                if (is_zero_host(Hm(i, i + 1))) {
                    i--;                 // happy breakdown
                    break_flag = true;
                    break;
                }
                for (int j = 1; j < i; j++) {
                    const double h1 = c1[j] * Hm(i, j) + c2[j] * Hm(i, j + 1);
                    const double h2 = -c2[j] * Hm(i, j) + c1[j] * Hm(i, j + 1);
                    Hm(i, j) = h1;
                    Hm(i, j + 1) = h2;
                }
                double tmp = sqrt(Hm(i, i) * Hm(i, i) + Hm(i, i + 1) * Hm(i, i + 1));
                if (is_zero_host(tmp))
                    tmp = 1.e-16;
                c1[i] = Hm(i, i) / tmp;
                c2[i] = Hm(i, i + 1) / tmp;
                sv[i + 1] = -c2[i] * sv[i];
                sv[i] = c1[i] * sv[i];
                Hm(i, i) = c1[i] * Hm(i, i) + c2[i] * Hm(i, i + 1);
                r_norm = fabs(sv[i + 1]);
                if (left) {
                    // false convergence: run on until the estimate has
                    // dropped by the ratio true / estimated residual
                    if (false_conv && r_norm * s0 <= tol)
                        false_conv = false;
                }
                else if (r_norm <= tol)
                    break;
            }
            for (int j = i; j > 0; j--) {
                yv[j] = sv[j] / Hm(j, j);
                for (int k = j - 1; k > 0; k--)
                    sv[k] -= Hm(j, k) * yv[j];
            }
            // flexible: x += sum_j y_j flv_j (x += z after a cycle of no step)
            if (flex) {
                if (i > 0) {
                    const double *cd = put_coef(0, &yv[1], i);
                    if (cd == nullptr)
                        return phg::set_error(PHG_HIP_E_HIP, "GMRES coefficient copy failed");
                    hipLaunchKernelGGL(k_gm_combo, dim3(g), dim3(BS), 0, s, n,
                                       (const double *)FL(1), i, (const double *)nullptr, 0,
                                       cd, 0.0, (double *)nullptr, xx,
                                       (const double *)nullptr, (double *)nullptr);
                }
                else
                    hipLaunchKernelGGL(k_gm_combo, dim3(g), dim3(BS), 0, s, n,
                                       (const double *)z, 1, (const double *)nullptr, 0,
                                       (const double *)nullptr, 1.0, (double *)nullptr, xx,
                                       (const double *)nullptr, (double *)nullptr);
            }
            // z = V y plus the augmentation part (y_{mk+j} for j <= the
            // number of augmentation vectors: values of an earlier cycle
            // where this one stopped before step mk + j, as in the
            // reference), kept as augmentation vector itr_out % augk; the
            // left variants add it to x, the plain right ones M^-1 of it
            if (!flex || aug) {
                double *xo = flex ? (double *)nullptr : xx;
                const double *pd = left ? (const double *)nullptr : d;
                double *zo = aug ? ZP(itr_out % augk) : (double *)nullptr;
                if (i > 0) {
                    const int top = i <= mk ? i : mk;
                    const int ak = i <= mk ? 0 : (itr_out < augk ? itr_out : augk);
                    std::copy(&yv[1], &yv[1] + top, cbuf.begin());
                    for (int j = 1; j <= ak; j++)
                        cbuf[(size_t)(top + j - 1)] = yv[j + mk];
                    const double *cd = put_coef(1, cbuf.data(), top + ak);
                    if (cd == nullptr)
                        return phg::set_error(PHG_HIP_E_HIP, "GMRES coefficient copy failed");
                    hipLaunchKernelGGL(k_gm_combo, dim3(g), dim3(BS), 0, s, n,
                                       (const double *)V(1), top, (const double *)ZP(0), ak,
                                       cd, 0.0, z, xo, pd, zo);
                }
                else
                    hipLaunchKernelGGL(k_gm_combo, dim3(g), dim3(BS), 0, s, n,
                                       (const double *)z, 1, (const double *)nullptr, 0,
                                       (const double *)nullptr, 1.0, z, xo, pd, zo);
            }
            PHG_HIP_CHECK(hipGetLastError());
            if (!left) {
                if (r_norm <= tol) {
                    res = r_norm;
                    break;
                }
                double rr;
                if ((rc = residual_pass(1, &rr, &zz, nullptr)))
                    return rc;
            }
            else {
                // true residual r = b - A x and z = M^-1 r for the next cycle
                false_conv = false;
                double rr;
                if ((rc = residual_pass(1, &rr, &zz, nullptr)))
                    return rc;
                if (r_norm <= tol) {
                    res = rr;
                    if (res <= tol || break_flag)
                        break;
                    if (i < rs) {
                        false_conv = true;
                        s0 = res / r_norm;
                    }
                }
            }
            itr_out++;
        }
    }
    if (perm)
        hipLaunchKernelGGL(k_scatter, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)ws->xp, x->d);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    if (residual)
        *residual = res;
    return nits;
}

// FP64 PCG in the numbering of the format
template <typename View>
int pcg64_impl(phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x,
               int32_t maxit, double rtol, double btol, double atol,
               double *residual)
{
    using F = FmtData<View>;
    hipStream_t s = phg::ctx()->stream;
    Workspace *ws = get_ws(A, 0);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "PCG workspace allocation failed");
    State h;
    memset(&h, 0, sizeof(h));
    h.rtol = rtol;
    h.btol = btol;
    h.atol = atol;
    h.maxit = maxit;
    PHG_HIP_CHECK(hipMemcpyAsync(ws->st, &h, sizeof(State),
                                 hipMemcpyHostToDevice, s));
    const int64_t n = A->n;
    int g = grid_for(n), gs = grid_spmv<View>(A, k_pcg_spmv<double, View>);
    const double *bb = b->d;
    double *xx = x->d;
    const int32_t *perm = F::perm(A);
    if (perm) {
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)b->d, ws->bp);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)x->d, ws->xp);
        bb = ws->bp;
        xx = ws->xp;
    }
    hipLaunchKernelGGL(k_pcg_init_norms, dim3(g), dim3(BS), 0, s, n, bb,
                       (const double *)xx, ws->seg(SEG_NORMS));
    hipLaunchKernelGGL((k_pcg_init_res<View>), dim3(gs), dim3(BS), 0, s, n,
                       view_of(A, (View *)nullptr), F::v64(A), F::d64(A), bb,
                       xx, ws->r, (const double *)ws->seg(SEG_NORMS), g,
                       ws->seg(SEG_RES), ws->st);
    PHG_HIP_CHECK(hipGetLastError());
    State fin;
    int rc;
    // the prologue ran on the SpMV grid: its partials count is gs
    if ((rc = run_pcg_loop<double, View>(A, F::v64(A), F::d64(A), xx, ws->r,
                                         ws->p, ws->Ap, ws, maxit, gs, 8, &fin)))
        return rc;
    if (perm)
        hipLaunchKernelGGL(k_scatter, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)ws->xp, x->d);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    // |b| = 0: solver-phg.c:497-501 returns before setting *res
    if (residual && !is_zero_host(fin.rb))
        *residual = fin.res;
    if (fin.done == ZERO_DIAG)
        return phg::set_error(PHG_HIP_E_ZERO_DIAG,
                              "zero diagonal in Jacobi preconditioner");
    return fin.nits;
}

// FP32 PCG under FP64 refinement in the numbering of the format
template <typename View>
int mixed_impl(phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x,
               int32_t maxit, double rtol, double btol, double atol,
               double inner_rtol, int32_t *inner_its, double *residual)
{
    using F = FmtData<View>;
    hipStream_t s = phg::ctx()->stream;
    Workspace *ws = get_ws(A, 1);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "workspace allocation failed");
    const int64_t n = A->n;
    int g = grid_for(n), gs = grid_spmv<View>(A, k_pcg_spmv<float, View>);
    double *hm = (double *)phg::ctx()->host_mailbox;
    double tol = 0.0, res = 0.0;
    int total_inner = 0, outer = 0, rc;
    const double *bb = b->d;
    double *xx = x->d;
    const int32_t *perm = F::perm(A);
    if (perm) {
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)b->d, ws->bp);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)x->d, ws->xp);
        bb = ws->bp;
        xx = ws->xp;
    }
    for (;;) {
        hipLaunchKernelGGL((k_ir_residual<View>), dim3(gs), dim3(BS), 0, s,
                           view_of(A, (View *)nullptr), F::v64(A), bb,
                           (const double *)xx, ws->r, ws->seg(SEG_IR));
        hipLaunchKernelGGL(k_ir_norms, dim3(1), dim3(BS), 0, s,
                           (const double *)ws->seg(SEG_IR), gs, ws->scal);
        PHG_HIP_CHECK(hipGetLastError());
        PHG_HIP_CHECK(hipMemcpyAsync(hm, ws->scal, 2 * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
        PHG_HIP_CHECK(hipStreamSynchronize(s));
        res = hm[0];
        const double rb = hm[1];
        if (outer == 0) {
            // get_tol on the FP64 residual (solver-phg.c:104)
            tol = res * rtol;
            if (tol < rb * btol)
                tol = rb * btol;
            if (tol < atol)
                tol = atol;
        }
        if (res <= tol || rb == 0.0 || outer >= maxit)
            break;
        State h;
        memset(&h, 0, sizeof(h));
        h.rtol = inner_rtol;
        h.rb = 1.0;
        h.maxit = 100000;     // each FP32 solve stops on inner_rtol
        PHG_HIP_CHECK(hipMemcpyAsync(ws->st, &h, sizeof(State),
                                     hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_ir_inner_init, dim3(g), dim3(BS), 0, s, n,
                           (const double *)ws->r, 1.0 / res, F::d32(A),
                           ws->r32, ws->x32, ws->seg(SEG_RES));
        PHG_HIP_CHECK(hipGetLastError());
        State fin;
        if ((rc = run_pcg_loop<float, View>(A, F::v32(A), F::d32(A), ws->x32,
                                            ws->r32, ws->p32, ws->Ap32, ws,
                                            h.maxit, g, 16, &fin)))
            return rc;
        if (fin.done == ZERO_DIAG)
            return phg::set_error(PHG_HIP_E_ZERO_DIAG,
                                  "zero diagonal in Jacobi preconditioner");
        total_inner += fin.nits;
        hipLaunchKernelGGL(k_ir_correct, dim3(g), dim3(BS), 0, s, n, xx,
                           (const float *)ws->x32, res);
        PHG_HIP_CHECK(hipGetLastError());
        outer++;
    }
    if (perm)
        hipLaunchKernelGGL(k_scatter, dim3(g), dim3(BS), 0, s, n, perm,
                           (const double *)ws->xp, x->d);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    if (inner_its)
        *inner_its = total_inner;
    if (residual)
        *residual = res;
    return outer;
}

// FP32 PCG with reliable FP64 residual updates, in pieces: the prologue
// (gather into the solver numbering, r = b - A x), one slot of launches
// (k_ru_p, k_ru_spmv, k_ru_update, k_ru_resid) and the end (fold x32 into x,
// scatter back).  Slots write the ping-pong state st[(k + 1) & 1].
template <typename View>
struct RuPlan {
    View mv;
    int g = 0, gs = 0;
    // FP32 SpMV through the matrix-free operator when eo is set; gs32 is the
    // number of partials the FP32 SpMV writes
    const EopData *eo = nullptr;
    int gs32 = 0;
    // the FP64 residual of UPDATE slots runs in k_ru_update (layouts in
    // the solver numbering) or in k_ru_resid (layouts with their own order)
    bool resid_in_update = true;
    const double *bb = nullptr;
    double *xx = nullptr;
    const int32_t *perm = nullptr;
    Workspace *ws = nullptr;
};

template <typename View>
int ru_prologue(phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x,
                int32_t maxit, double rtol, double btol, double atol,
                double delta, hipStream_t s, RuPlan<View> *pl, bool capture = false)
{
    using F = FmtData<View>;
    Workspace *ws = get_ws(A, 1);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "workspace allocation failed");
    const int64_t n = A->n;
    pl->ws = ws;
    pl->g = grid_for(n);
    pl->gs = grid_spmv<View>(A, k_ru_spmv<View>);
    pl->mv = view_of(A, (View *)nullptr);
    pl->gs32 = pl->gs;
    // the operator works in the solver numbering, as every layout but SELL
    const bool solver_numbering = F::perm(A) == A->q_perm;
    pl->resid_in_update = solver_numbering && !std::is_same<View, PhgPermView>::value;
    if (A->eop != nullptr && A->eop->on && solver_numbering) {
        pl->eo = A->eop;
        pl->gs32 = A->eop->grid_b;
    }
    State h;
    memset(&h, 0, sizeof(h));
    h.rtol = rtol;
    h.btol = btol;
    h.atol = atol;
    h.maxit = maxit;
    h.delta = delta;
    h.mode = RU_START;
    if (capture) {
        // under graph capture: the state is set by a kernel, no host copy
        hipLaunchKernelGGL(k_state_set, dim3(1), dim3(64), 0, s, ws->st, h);
    }
    else {
        PHG_HIP_CHECK(hipMemcpyAsync(ws->st, &h, sizeof(State),
                                     hipMemcpyHostToDevice, s));
        PHG_HIP_CHECK(hipStreamSynchronize(s));    // h lives on this stack
    }
    pl->bb = b->d;
    pl->xx = x->d;
    pl->perm = F::perm(A);
    const int g = pl->g;
    if (pl->perm) {
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, pl->perm,
                           (const double *)b->d, ws->bp);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(BS), 0, s, n, pl->perm,
                           (const double *)x->d, ws->xp);
        pl->bb = ws->bp;
        pl->xx = ws->xp;
    }
    hipLaunchKernelGGL(k_pcg_init_norms, dim3(g), dim3(BS), 0, s, n, pl->bb,
                       (const double *)pl->xx, ws->seg(SEG_NORMS));
    hipLaunchKernelGGL((k_ru_init<View>), dim3(pl->gs), dim3(BS), 0, s, n,
                       pl->mv, F::v64(A), F::d64(A), pl->bb, pl->xx, ws->r32,
                       ws->x32, (const double *)ws->seg(SEG_NORMS), g,
                       ws->seg(SEG_RES), ws->st);
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

template <typename View>
void ru_slot(const phg_hip_mat *A, const RuPlan<View> &pl, int64_t k,
             hipStream_t s, bool prof)
{
    using F = FmtData<View>;
    Workspace *ws = pl.ws;
    const int64_t n = A->n;
    const int g = pl.g, gs = pl.gs;
    double *pres = ws->seg(SEG_RES), *ps = ws->seg(SEG_S);
    State *sin = ws->st + (k & 1), *sout = ws->st + ((k + 1) & 1);
    if (prof) prof_mark(KP_P32, s);
    hipLaunchKernelGGL(k_ru_p, dim3(g), dim3(BS), 0, s, n,
                       (const float *)ws->r32, F::d32(A), ws->p32, pl.xx,
                       ws->x32, (const double *)pres, g,
                       pl.resid_in_update ? g : gs, gs, (const State *)sin, sout);
    if (prof) { prof_mark(KP_P32, s); prof_mark(KP_SPMV32, s); }
    // the operator's shared rows are summed in k_ru_update when their codes
    // exist, else by the fix-up kernel
    if (pl.eo != nullptr)
        phg::eop_apply(pl.eo, ws->p32, F::d32(A), ws->Ap32, ps, &sout->done, &sout->mode,
                       RU_ITER, s, 0, pl.eo->fxc == nullptr);
    else
        hipLaunchKernelGGL((k_ru_spmv<View>), dim3(gs), dim3(BS), 0, s, pl.mv,
                           F::v32(A), (const float *)ws->p32, ws->Ap32, ps,
                           (const State *)sout);
    if (prof) { prof_mark(KP_SPMV32, s); prof_mark(KP_UPD32, s); }
    hipLaunchKernelGGL(k_ru_update, dim3(g), dim3(BS), 0, s, n,
                       ws->x32, (const float *)ws->p32, ws->r32,
                       (const float *)ws->Ap32, F::d32(A),
                       (const double *)ps, pl.gs32, pres, sout,
                       (const int32_t *)A->q_rowptr, (const int32_t *)A->q_cols,
                       (const double *)A->q_vals, (const int32_t *)A->q_rblk,
                       A->q_nrblk, (const double *)A->q_diag,
                       pl.bb, (const double *)pl.xx, pl.resid_in_update ? 1 : 0,
                       pl.eo ? (const uint32_t *)pl.eo->fxc : nullptr,
                       pl.eo ? (const float *)pl.eo->part : nullptr);
    if (prof) prof_mark(KP_UPD32, s);
    if (!pl.resid_in_update)
        hipLaunchKernelGGL((k_ru_resid<View>), dim3(gs), dim3(BS), 0, s,
                           pl.mv, F::v64(A), F::d64(A), pl.bb,
                           (const double *)pl.xx, ws->r32, pres,
                           (const State *)sout);
}

template <typename View>
void ru_finish(const phg_hip_mat *A, const RuPlan<View> &pl, phg_hip_vec *x,
               hipStream_t s)
{
    const int64_t n = A->n;
    hipLaunchKernelGGL(k_ru_fold, dim3(pl.g), dim3(BS), 0, s, n, pl.xx,
                       (const float *)pl.ws->x32);
    if (pl.perm)
        hipLaunchKernelGGL(k_scatter, dim3(pl.g), dim3(BS), 0, s, n, pl.perm,
                           (const double *)pl.ws->xp, x->d);
}

// FP32 PCG with reliable FP64 residual updates, in the numbering of the
// format; returns the number of FP32 iterations
template <typename View>
int reliable_impl(phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x,
                  int32_t maxit, double rtol, double btol, double atol,
                  double delta, int32_t *updates, double *residual)
{
    hipStream_t s = phg::ctx()->stream;
    RuPlan<View> pl;
    int rc;
    if ((rc = ru_prologue<View>(A, b, x, maxit, rtol, btol, atol, delta, s, &pl)))
        return rc;
    Workspace *ws = pl.ws;
    State *hm = (State *)phg::ctx()->host_mailbox;
    const bool prof = g_prof.on;
    int64_t k = 0;
    int chunk = prof ? 1 : 16;
    for (;;) {
        for (int i = 0; i < chunk; i++, k++)
            ru_slot<View>(A, pl, k, s, prof);
        PHG_HIP_CHECK(hipGetLastError());
        if ((rc = read_state(ws->st + (k & 1), hm, s)))
            return rc;
        if (prof) {
            // keep the timing of the kernels that did the work of the slot
            const int drop[3] = {hm->mode == RU_ITER ? -1 : KP_SPMV32,
                                 hm->mode == RU_ITER ? -1 : KP_UPD32, -1};
            for (int d : drop)
                if (d >= 0 && !hm->done)
                    g_prof.used[d] -= 2;
            if (hm->done)
                for (int d : {KP_P32, KP_SPMV32, KP_UPD32})
                    g_prof.used[d] -= 2;
        }
        if (hm->done != RUNNING)
            break;
        if (k > 2 * (int64_t)maxit + 4096)
            return phg::set_error(PHG_HIP_E_HIP, "PCG state did not advance");
        if (!prof)
            chunk = std::min(chunk * 2, 64);
    }
    if (prof)
        prof_collect();
    State fin = *hm;
    ru_finish<View>(A, pl, x, s);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    if (updates)
        *updates = fin.updates;
    if (residual)
        *residual = fin.res;
    if (fin.done == ZERO_DIAG)
        return phg::set_error(PHG_HIP_E_ZERO_DIAG,
                              "zero diagonal in Jacobi preconditioner");
    return fin.nits;
}

// One matrix of a batch: its type-erased slot launcher and state.
struct MultiItem {
    std::function<void(int64_t, hipStream_t)> slot;
    std::function<void(hipStream_t)> finish;
    State *st = nullptr;
};

template <typename View>
int multi_prepare(phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x,
                  int32_t maxit, double rtol, double btol, double atol,
                  double delta, hipStream_t s, MultiItem *it)
{
    auto pl = std::make_shared<RuPlan<View>>();
    int rc;
    if ((rc = ru_prologue<View>(A, b, x, maxit, rtol, btol, atol, delta, s, pl.get())))
        return rc;
    it->slot = [A, pl](int64_t k, hipStream_t st) {
        ru_slot<View>(A, *pl, k, st, false);
    };
    it->finish = [A, pl, x](hipStream_t st) { ru_finish<View>(A, *pl, x, st); };
    it->st = pl->ws->st;
    return PHG_HIP_OK;
}

// time `reps` launches of the PCG SpMV kernel of one format and precision
template <typename T, typename View>
int time_spmv(phg_hip_mat *A, const T *va, int reps, double *us)
{
    hipStream_t s = phg::ctx()->stream;
    Workspace *ws = get_ws(A, 1);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "workspace allocation failed");
    State h;
    memset(&h, 0, sizeof(h));
    PHG_HIP_CHECK(hipMemcpyAsync(ws->st + 2, &h, sizeof(State),
                                 hipMemcpyHostToDevice, s));
    T *p = (T *)(sizeof(T) == 8 ? (void *)ws->p : (void *)ws->p32);
    T *Ap = (T *)(sizeof(T) == 8 ? (void *)ws->Ap : (void *)ws->Ap32);
    PHG_HIP_CHECK(hipMemsetAsync(p, 0, (size_t)A->n * sizeof(T), s));
    int gs = grid_spmv<View>(A, k_pcg_spmv<T, View>);
    hipEvent_t e0, e1;
    PHG_HIP_CHECK(hipEventCreate(&e0));
    PHG_HIP_CHECK(hipEventCreate(&e1));
    for (int i = 0; i < reps + 2; i++) {
        if (i == 2)
            PHG_HIP_CHECK(hipEventRecord(e0, s));
        hipLaunchKernelGGL((k_pcg_spmv<T, View>), dim3(gs), dim3(BS), 0, s,
                           view_of(A, (View *)nullptr), va, (const T *)p, Ap,
                           ws->seg(SEG_S), (const State *)(ws->st + 2));
    }
    hipError_t err = hipGetLastError();
    float ms = 0.f;
    if (err == hipSuccess)
        err = hipEventRecord(e1, s);
    if (err == hipSuccess)
        err = hipEventSynchronize(e1);
    if (err == hipSuccess)
        err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess)
        return phg::set_error(PHG_HIP_E_HIP, "SpMV timing (grid %d, %lld units): %s",
                              gs, (long long)units_of(A, (View *)nullptr),
                              hipGetErrorString(err));
    *us = 1e3 * ms / reps;
    return PHG_HIP_OK;
}

}  // namespace

namespace phg {

// Greedy split of the rows into runs of consecutive rows holding at most
// cap entries and rcap rows; a longer row forms a block of its own.  The
// greedy runs on its own in chunks of SPLIT_CHUNK rows (a block never spans
// two chunks), one work-item per chunk: a counting pass, an exclusive scan
// of the counts, and a pass writing each chunk's block ends at its offset.
// A row's sum in the kernels does not depend on the blocking.
constexpr int64_t SPLIT_CHUNK = 1024;

template <typename Emit>
__device__ inline void split_chunk(const int32_t *rp, int64_t r0, int64_t r1, int64_t cap,
                                   int64_t rcap, Emit emit)
{
    int64_t cur_nnz = 0, cur_rows = 0;
    for (int64_t r = r0; r < r1; r++) {
        const int64_t len = rp[r + 1] - rp[r];
        if (len > cap) {
            if (cur_rows > 0)
                emit((int32_t)r);
            emit((int32_t)(r + 1));
            cur_nnz = cur_rows = 0;
            continue;
        }
        if (cur_rows > 0 && (cur_nnz + len > cap || cur_rows == rcap)) {
            emit((int32_t)r);
            cur_nnz = cur_rows = 0;
        }
        cur_nnz += len;
        cur_rows++;
    }
    if (cur_rows > 0)
        emit((int32_t)r1);
}

__global__ void k_split_count(const int32_t *rp, int64_t n, int64_t cap, int64_t rcap,
                              int32_t *cnt)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r0 = c * SPLIT_CHUNK;
    if (r0 >= n)
        return;
    int32_t k = 0;
    split_chunk(rp, r0, min(n, r0 + SPLIT_CHUNK), cap, rcap, [&](int32_t) { k++; });
    cnt[c] = k;
}

// exclusive scan of cnt[0..m) into off[0..m], one work-group of BS
__global__ void __launch_bounds__(BS) k_split_scan(const int32_t *cnt, int64_t m,
                                                   int32_t *off)
{
    __shared__ int32_t sh[BS];
    int32_t carry = 0;
    for (int64_t base = 0; base < m; base += BS) {
        const int64_t i = base + threadIdx.x;
        int32_t v = i < m ? cnt[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < BS; o <<= 1) {
            const int32_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < m)
            off[i] = carry + sh[threadIdx.x] - v;
        const int32_t tot = sh[BS - 1];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0)
        off[m] = carry;
}

__global__ void k_split_fill(const int32_t *rp, int64_t n, int64_t cap, int64_t rcap,
                             const int32_t *off, int32_t *blk)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r0 = c * SPLIT_CHUNK;
    if (c == 0)
        blk[0] = 0;
    if (r0 >= n)
        return;
    int32_t *o = blk + 1 + off[c];
    split_chunk(rp, r0, min(n, r0 + SPLIT_CHUNK), cap, rcap, [&](int32_t e) { *o++ = e; });
}

int split_row_blocks(const int32_t *d_rowptr, int64_t n, int64_t nnz,
                     int64_t cap, int64_t rcap, int32_t **d, int64_t *nb)
{
    (void)nnz;
    hipStream_t s = ctx()->stream;
    const int64_t m = (n + SPLIT_CHUNK - 1) / SPLIT_CHUNK;
    int32_t *cnt = nullptr, *off = nullptr, total = 0;
    int rc;
    dfree(*d);
    *d = nullptr;
    // at most two block ends per row, plus the leading 0
    if ((rc = dmalloc((void **)&cnt, (size_t)(m + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&off, (size_t)(m + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)d, (size_t)(2 * n + 2) * sizeof(int32_t)))) {
        dfree(cnt);
        dfree(off);
        dfree(*d);
        *d = nullptr;
        return rc;
    }
    const int g = (int)std::max<int64_t>(1, (m + 255) / 256);
    hipLaunchKernelGGL(k_split_count, dim3(g), dim3(256), 0, s, d_rowptr, n, cap, rcap, cnt);
    hipLaunchKernelGGL(k_split_scan, dim3(1), dim3(BS), 0, s, (const int32_t *)cnt, m, off);
    hipLaunchKernelGGL(k_split_fill, dim3(g), dim3(256), 0, s, d_rowptr, n, cap, rcap,
                       (const int32_t *)off, *d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        rc = download(&total, off + m, sizeof(int32_t));
    else
        rc = hip_fail(e, "row-block split", __FILE__, __LINE__);
    dfree(cnt);
    dfree(off);
    if (rc) {
        dfree(*d);
        *d = nullptr;
        return rc;
    }
    *nb = total;
    return PHG_HIP_OK;
}

int build_row_blocks(phg_hip_mat *A)
{
    return split_row_blocks(A->d_rowptr, A->n, A->nnz, NNZ_CAP, ROW_CAP,
                            &A->d_rblk, &A->nrblk);
}

void free_sell(phg_hip_mat *A)
{
    dfree(A->d_perm); dfree(A->d_iperm); dfree(A->d_sptr); dfree(A->d_swidth);
    dfree(A->d_scols); dfree(A->d_svals); dfree(A->d_svals32);
    dfree(A->d_sdiag); dfree(A->d_sdiag32);
    A->d_perm = A->d_iperm = A->d_swidth = A->d_scols = nullptr;
    A->d_sptr = nullptr;
    A->d_svals = A->d_sdiag = nullptr;
    A->d_svals32 = A->d_sdiag32 = nullptr;
    A->nslice = A->sell_nnz = 0;
}

constexpr int64_t SELL_WINDOW = 4096;   // rows sorted together (64 slices)

int build_sell(phg_hip_mat *A)
{
    free_sell(A);
    const int64_t n = A->n;
    std::vector<int32_t> rp((size_t)n + 1);
    int rc;
    if ((rc = download(rp.data(), A->q_rowptr, rp.size() * sizeof(int32_t))))
        return rc;
    std::vector<int32_t> qperm;
    if (A->q_perm) {
        qperm.resize((size_t)n);
        if ((rc = download(qperm.data(), A->q_perm, qperm.size() * sizeof(int32_t))))
            return rc;
    }
    // stable sort by decreasing length inside each window: rows of equal
    // length keep their order, so neighbouring rows stay neighbours
    std::vector<int32_t> perm((size_t)n), iperm((size_t)n);
    std::vector<int32_t> order;
    for (int64_t w0 = 0; w0 < n; w0 += SELL_WINDOW) {
        int64_t w1 = std::min<int64_t>(w0 + SELL_WINDOW, n);
        order.resize((size_t)(w1 - w0));
        for (int64_t i = w0; i < w1; i++)
            order[(size_t)(i - w0)] = (int32_t)i;
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return rp[(size_t)a + 1] - rp[(size_t)a] > rp[(size_t)b + 1] - rp[(size_t)b];
        });
        for (int64_t i = w0; i < w1; i++)
            perm[(size_t)i] = order[(size_t)(i - w0)];
    }
    // perm/iperm relate SELL rows to solver-CSR rows here; the stored
    // permutation maps SELL rows to PHG rows
    for (int64_t i = 0; i < n; i++)
        iperm[(size_t)perm[(size_t)i]] = (int32_t)i;
    std::vector<int32_t> to_phg(perm);
    if (A->q_perm)
        for (int64_t i = 0; i < n; i++)
            to_phg[(size_t)i] = qperm[(size_t)perm[(size_t)i]];
    const int64_t nslice = (n + 63) / 64;
    std::vector<int64_t> sptr((size_t)nslice + 1);
    std::vector<int32_t> width((size_t)nslice);
    sptr[0] = 0;
    for (int64_t sl = 0; sl < nslice; sl++) {
        int32_t w = 0;
        for (int64_t i = sl * 64; i < std::min<int64_t>(sl * 64 + 64, n); i++) {
            int32_t r = perm[(size_t)i];
            w = std::max(w, rp[(size_t)r + 1] - rp[(size_t)r]);
        }
        width[(size_t)sl] = w;
        sptr[(size_t)sl + 1] = sptr[(size_t)sl] + 64 * (int64_t)w;
    }
    const int64_t snnz = sptr[(size_t)nslice];
    A->nslice = nslice;
    A->sell_nnz = snnz;
    if ((rc = dmalloc((void **)&A->d_perm, (size_t)n * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&A->d_iperm, (size_t)n * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&A->d_sptr, sptr.size() * sizeof(int64_t))) ||
        (rc = dmalloc((void **)&A->d_swidth, width.size() * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&A->d_scols, (size_t)(snnz + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&A->d_svals, (size_t)(snnz + 1) * sizeof(double))) ||
        (rc = dmalloc((void **)&A->d_sdiag, (size_t)n * sizeof(double))) ||
        (rc = upload(A->d_perm, perm.data(), perm.size() * sizeof(int32_t))) ||
        (rc = upload(A->d_iperm, iperm.data(), iperm.size() * sizeof(int32_t))) ||
        (rc = upload(A->d_sptr, sptr.data(), sptr.size() * sizeof(int64_t))) ||
        (rc = upload(A->d_swidth, width.data(), width.size() * sizeof(int32_t)))) {
        free_sell(A);
        return rc;
    }
    hipStream_t s = ctx()->stream;
    hipLaunchKernelGGL(k_sell_fill, dim3((unsigned)((nslice * 64 + 255) / 256)),
                       dim3(256), 0, s, n, nslice, (const int32_t *)A->q_rowptr,
                       (const int32_t *)A->q_cols, (const double *)A->q_vals,
                       (const double *)A->q_diag, (const int32_t *)A->d_perm,
                       (const int32_t *)A->d_iperm, (const int64_t *)A->d_sptr,
                       (const int32_t *)A->d_swidth, A->d_scols, A->d_svals,
                       A->d_sdiag);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    // the solvers gather with the SELL -> PHG map
    if ((rc = upload(A->d_perm, to_phg.data(), to_phg.size() * sizeof(int32_t))))
        return rc;
    return PHG_HIP_OK;
}

int ensure_fp32_copy(phg_hip_mat *A)
{
    hipStream_t s = ctx()->stream;
    int rc;
    if (A->q_vals32 == nullptr) {
        if ((rc = dmalloc((void **)&A->q_vals32, (size_t)(A->nnz + 1) * sizeof(float))) ||
            (rc = dmalloc((void **)&A->q_diag32, (size_t)A->n * sizeof(float))))
            return rc;
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->nnz)), dim3(BS), 0, s,
                           A->nnz, (const double *)A->q_vals, A->q_vals32);
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->n)), dim3(BS), 0, s, A->n,
                           (const double *)A->q_diag, A->q_diag32);
    }
    if (A->d_svals != nullptr && A->d_svals32 == nullptr) {
        if ((rc = dmalloc((void **)&A->d_svals32, (size_t)(A->sell_nnz + 1) * sizeof(float))) ||
            (rc = dmalloc((void **)&A->d_sdiag32, (size_t)A->n * sizeof(float))))
            return rc;
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->sell_nnz)), dim3(BS), 0, s,
                           A->sell_nnz, (const double *)A->d_svals, A->d_svals32);
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->n)), dim3(BS), 0, s, A->n,
                           (const double *)A->d_sdiag, A->d_sdiag32);
    }
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    return PHG_HIP_OK;
}

void free_workspace(phg_hip_mat *A)
{
    Workspace *w = (Workspace *)A->ws;
    if (w) {
        release_ws(w);
        delete w;
        A->ws = nullptr;
    }
}

int spmv64(const phg_hip_mat *A, const double *x, double *y, double alpha,
           double beta, hipStream_t s)
{
    // the row blocks of the PHG-order CSR are built on the first product
    if (A->d_rblk == nullptr && A->n > 0) {
        int rc = build_row_blocks(const_cast<phg_hip_mat *>(A));
        if (rc)
            return rc;
    }
    if (A->nrblk == 0)
        return PHG_HIP_OK;
    hipLaunchKernelGGL(k_spmv64, dim3(grid_spmv<CsrView>(A, k_spmv64)), dim3(BS), 0, s,
                       (const int32_t *)A->d_rowptr, (const int32_t *)A->d_cols,
                       (const double *)A->d_vals, (const int32_t *)A->d_rblk,
                       A->nrblk, x, y, alpha, beta);
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

}  // namespace phg

namespace {
int ensure_format(phg_hip_mat *A, int fmt, int need32)
{
    int rc;
    if ((rc = phg::ensure_solver_copy(A)))
        return rc;
    // the 4096-entry and wavefront row blocks of the solver copy are split
    // when a layout first uses them
    if ((fmt == FMT_B4K || fmt == FMT_B4K_PAR) && A->q_rblk1k == nullptr &&
        (rc = phg::split_row_blocks(A->q_rowptr, A->n, A->nnz, phg::SPMV_LARGE_NNZ,
                                    phg::SPMV_BLOCK_ROWS, &A->q_rblk1k, &A->q_nrblk1k)))
        return rc;
    if (fmt == FMT_WAVE && A->q_wblk == nullptr &&
        (rc = phg::split_row_blocks(A->q_rowptr, A->n, A->nnz, phg::SPMV_WAVE_NNZ,
                                    phg::SPMV_WAVE_ROWS, &A->q_wblk, &A->q_nwblk)))
        return rc;
    if (fmt == FMT_SELL && A->d_svals == nullptr && (rc = phg::build_sell(A)))
        return rc;
    if (fmt == FMT_DICT || fmt == FMT_DICT_NP) {
        if ((rc = phg::build_dict(A)))
            return rc;
        if (A->cd.state != 1)
            return phg::set_error(PHG_HIP_E_ARG,
                                  "matrix has no dictionary-compressed layout");
    }
    if (fmt == FMT_XS || fmt == FMT_XS_PF) {
        if ((rc = phg::build_xstage(A)))
            return rc;
        if (A->xs.state != 1)
            return phg::set_error(PHG_HIP_E_ARG,
                                  "matrix has no staged-line layout");
    }
    if (need32 && (rc = phg::ensure_fp32_copy(A)))
        return rc;
    return PHG_HIP_OK;
}
}  // namespace

//================================================================= C ABI

extern "C" int phg_hip_matvec(double alpha, const phg_hip_mat *A,
                              const phg_hip_vec *x, double beta,
                              phg_hip_vec *y)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr || x == nullptr || y == nullptr || x->n != A->n ||
        y->n != A->n)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_matvec: size mismatch");
    if (x->d == y->d)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_matvec: x and y alias");
    return phg::spmv64(A, x->d, y->d, alpha, beta, phg::ctx()->stream);
}

extern "C" int phg_hip_vec_dot(const phg_hip_vec *x, const phg_hip_vec *y,
                               double *res)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (!x || !y || !res || x->n != y->n)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_vec_dot: bad args");
    hipStream_t s = phg::ctx()->stream;
    double *part = nullptr;
    if ((rc = phg::dmalloc((void **)&part, (MAX_GRID + 1) * sizeof(double))))
        return rc;
    int g = grid_for(x->n);
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(BS), 0, s, x->n,
                       (const double *)x->d, (const double *)y->d, part);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BS), 0, s,
                       (const double *)part, g, part + MAX_GRID);
    PHG_HIP_CHECK(hipGetLastError());
    rc = phg::download(res, part + MAX_GRID, sizeof(double));
    phg::dfree(part);
    return rc;
}

extern "C" int phg_hip_pc_jacobi(const phg_hip_mat *A, const phg_hip_vec *b,
                                 phg_hip_vec *x)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (!A || !b || !x || b->n != A->n || x->n != A->n)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_pc_jacobi: bad args");
    if (!A->eliminated || A->d_diag == nullptr)
        return phg::set_error(PHG_HIP_E_ARG,
                              "phg_hip_pc_jacobi: matrix not eliminated (no diagonal)");
    hipStream_t s = phg::ctx()->stream;
    double *part = nullptr;
    if ((rc = phg::dmalloc((void **)&part, (MAX_GRID + 1) * sizeof(double))))
        return rc;
    int g = grid_for(A->n);
    hipLaunchKernelGGL(k_jacobi, dim3(g), dim3(BS), 0, s, A->n, (const double *)A->d_diag,
                       (const double *)b->d, x->d, part);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BS), 0, s,
                       (const double *)part, g, part + MAX_GRID);
    PHG_HIP_CHECK(hipGetLastError());
    double nzero = 0.0;
    rc = phg::download(&nzero, part + MAX_GRID, sizeof(double));
    phg::dfree(part);
    if (rc)
        return rc;
    if (nzero != 0.0)
        return phg::set_error(PHG_HIP_E_ZERO_DIAG,
                              "zero diagonal in Jacobi preconditioner (%.0f rows)", nzero);
    return PHG_HIP_OK;
}

extern "C" int phg_hip_vec_norm2(const phg_hip_vec *x, double *res)
{
    int rc = phg_hip_vec_dot(x, x, res);
    if (rc == PHG_HIP_OK)
        *res = sqrt(*res);
    return rc;
}

extern "C" int phg_hip_vec_axpby(double a, const phg_hip_vec *x, double b,
                                 phg_hip_vec *y)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (!x || !y || x->n != y->n)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_vec_axpby: bad args");
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(y->n)), dim3(BS), 0,
                       phg::ctx()->stream, y->n, a, (const double *)x->d, b,
                       y->d);
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

extern "C" int phg_hip_vec_copy(const phg_hip_vec *src, phg_hip_vec *dst)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (!src || !dst || src->n != dst->n)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_vec_copy: bad args");
    PHG_HIP_CHECK(hipMemcpyAsync(dst->d, src->d, (size_t)src->n * sizeof(double),
                                 hipMemcpyDeviceToDevice, phg::ctx()->stream));
    return PHG_HIP_OK;
}

extern "C" int phg_hip_pcg(const phg_hip_mat *A_, const phg_hip_vec *b,
                           phg_hip_vec *x, int32_t maxit, double rtol,
                           double btol, double atol, double *residual)
{
    int rc;
    if ((rc = check_solve_args(A_, b, x)))
        return rc;
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    if ((rc = ensure_format(A, A->fmt64, 0)))
        return rc;
    return with_format(A->fmt64, [&](auto tag) {
        using View = decltype(tag);
        return pcg64_impl<View>(A, b, x, maxit, rtol, btol, atol, residual);
    });
}

extern "C" int phg_hip_bicgstab(const phg_hip_mat *A_, const phg_hip_vec *b,
                                phg_hip_vec *x, int32_t maxit, double rtol,
                                double btol, double atol, double *residual)
{
    int rc;
    if ((rc = check_solve_args(A_, b, x)))
        return rc;
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    if ((rc = ensure_format(A, A->fmt64, 0)))
        return rc;
    return with_format(A->fmt64, [&](auto tag) {
        using View = decltype(tag);
        return bicgstab_impl<View>(A, b, x, maxit, rtol, btol, atol, residual);
    });
}

extern "C" int phg_hip_gmres_type(const phg_hip_mat *A_, const phg_hip_vec *b,
                                  phg_hip_vec *x, int32_t gmres_type, int32_t restart,
                                  int32_t augk, int32_t maxit, double rtol, double btol,
                                  double atol, double *residual)
{
    int rc;
    if ((rc = check_solve_args(A_, b, x)))
        return rc;
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    if ((rc = ensure_format(A, A->fmt64, 0)))
        return rc;
    return with_format(A->fmt64, [&](auto tag) {
        using View = decltype(tag);
        return gmres_impl<View>(A, b, x, gmres_type, restart, augk, maxit, rtol, btol, atol,
                                residual);
    });
}

extern "C" int phg_hip_gmres(const phg_hip_mat *A, const phg_hip_vec *b,
                             phg_hip_vec *x, int32_t restart, int32_t augk,
                             int32_t maxit, double rtol, double btol, double atol,
                             double *residual)
{
    return phg_hip_gmres_type(A, b, x, PHG_HIP_FLGMRES_R, restart, augk, maxit, rtol, btol,
                              atol, residual);
}

extern "C" int phg_hip_pcg_mixed(const phg_hip_mat *A_, const phg_hip_vec *b,
                                 phg_hip_vec *x, int32_t maxit, double rtol,
                                 double btol, double atol, double inner_rtol,
                                 int32_t *inner_its, double *residual)
{
    int rc;
    if ((rc = check_solve_args(A_, b, x)))
        return rc;
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    if ((rc = ensure_format(A, A->fmt32, 1)))
        return rc;
    return with_format(A->fmt32, [&](auto tag) {
        using View = decltype(tag);
        return mixed_impl<View>(A, b, x, maxit, rtol, btol, atol, inner_rtol,
                                inner_its, residual);
    });
}

extern "C" int phg_hip_pcg_reliable(const phg_hip_mat *A_,
                                    const phg_hip_vec *b, phg_hip_vec *x,
                                    int32_t maxit, double rtol, double btol,
                                    double atol, double delta,
                                    int32_t *updates, double *residual)
{
    int rc;
    if ((rc = check_solve_args(A_, b, x)))
        return rc;
    if (!(delta > 0.0 && delta < 1.0))
        return phg::set_error(PHG_HIP_E_ARG, "delta must lie in (0, 1)");
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    if ((rc = ensure_format(A, A->fmt32, 1)))
        return rc;
    return with_format(A->fmt32, [&](auto tag) {
        using View = decltype(tag);
        return reliable_impl<View>(A, b, x, maxit, rtol, btol, atol, delta,
                                   updates, residual);
    });
}

extern "C" int phg_hip_pcg_multi(int32_t nmat, phg_hip_mat *const *Av,
                                 const phg_hip_vec *const *b,
                                 phg_hip_vec *const *x, int32_t maxit,
                                 double rtol, double btol, double atol,
                                 double delta, int32_t use_graph,
                                 int32_t *nits, double *residual)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (nmat < 1 || nmat > MULTI_MAX || Av == nullptr || b == nullptr ||
        x == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_pcg_multi: 1 <= nmat <= %d",
                              MULTI_MAX);
    if (!(delta > 0.0 && delta < 1.0))
        return phg::set_error(PHG_HIP_E_ARG, "delta must lie in (0, 1)");
    hipStream_t s = phg::ctx()->stream;
    std::vector<MultiItem> items((size_t)nmat);
    for (int i = 0; i < nmat; i++) {
        if ((rc = check_solve_args(Av[i], b[i], x[i])))
            return rc;
        phg_hip_mat *A = Av[i];
        for (int j = 0; j < i; j++)
            if (Av[j] == A)
                return phg::set_error(PHG_HIP_E_ARG,
                                      "phg_hip_pcg_multi: a matrix appears twice");
        if ((rc = ensure_format(A, A->fmt32, 1)))
            return rc;
        rc = with_format(A->fmt32, [&](auto tag) {
            using View = decltype(tag);
            return multi_prepare<View>(A, b[i], x[i], maxit, rtol, btol, atol,
                                       delta, s, &items[(size_t)i]);
        });
        if (rc)
            return rc;
    }
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    // one stream per matrix (up to the pool size), forked from and joined to
    // the library stream
    hipStream_t *pool = nullptr;
    int npool = 0;
    if ((rc = phg::aux_streams(&pool, &npool)))
        return rc;
    const int CH = 16;          // slots per replay (even: the state parity holds)
    State *hm = (State *)phg::ctx()->host_mailbox;
    static_assert(sizeof(State) * MULTI_MAX <= 4096, "mailbox size");
    hipEvent_t fork, join[MULTI_MAX];
    PHG_HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (int i = 0; i < nmat; i++)
        PHG_HIP_CHECK(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
    auto enqueue = [&](hipStream_t main) -> hipError_t {
        hipError_t e;
        if ((e = hipEventRecord(fork, main)) != hipSuccess)
            return e;
        for (int i = 0; i < nmat; i++) {
            hipStream_t si = pool[i % npool];
            if ((e = hipStreamWaitEvent(si, fork, 0)) != hipSuccess)
                return e;
            for (int k = 0; k < CH; k++)
                items[(size_t)i].slot(k, si);
            if ((e = hipMemcpyAsync(hm + i, items[(size_t)i].st + (CH & 1),
                                    sizeof(State), hipMemcpyDeviceToHost, si)) !=
                hipSuccess)
                return e;
            if ((e = hipEventRecord(join[i], si)) != hipSuccess ||
                (e = hipStreamWaitEvent(main, join[i], 0)) != hipSuccess)
                return e;
        }
        return hipGetLastError();
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&]() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        (void)hipEventDestroy(fork);
        for (int i = 0; i < nmat; i++)
            (void)hipEventDestroy(join[i]);
    };
    hipError_t e = hipSuccess;
    if (use_graph) {
        if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeGlobal)) == hipSuccess) {
            hipError_t e1 = enqueue(s);
            e = hipStreamEndCapture(s, &graph);
            if (e1 != hipSuccess)
                e = e1;
        }
        if (e == hipSuccess)
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            cleanup();
            return phg::hip_fail(e, "graph capture of the batched solve", __FILE__,
                                 __LINE__);
        }
    }
    int64_t replays = 0;
    for (;;) {
        e = use_graph ? hipGraphLaunch(exec, s) : enqueue(s);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            cleanup();
            return phg::hip_fail(e, "batched solve", __FILE__, __LINE__);
        }
        replays++;
        bool all = true;
        for (int i = 0; i < nmat; i++)
            all = all && hm[i].done != RUNNING;
        if (all)
            break;
        if (replays * CH > 2 * (int64_t)maxit + 4096) {
            cleanup();
            return phg::set_error(PHG_HIP_E_HIP, "batched solve did not finish");
        }
    }
    std::vector<State> fin(hm, hm + nmat);
    cleanup();
    for (int i = 0; i < nmat; i++)
        items[(size_t)i].finish(s);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    int32_t mx = 0;
    for (int i = 0; i < nmat; i++) {
        if (fin[(size_t)i].done == ZERO_DIAG)
            return phg::set_error(PHG_HIP_E_ZERO_DIAG,
                                  "zero diagonal in Jacobi preconditioner");
        if (nits)
            nits[i] = fin[(size_t)i].nits;
        if (residual)
            residual[i] = fin[(size_t)i].res;
        mx = std::max(mx, fin[(size_t)i].nits);
    }
    return mx;
}

//=============================================================== batches
// Several problems assembled and solved in captured HIP graphs.  Per system
// (a stream of the pool each) the first graph holds the whole assembly on
// the device -- element pass, PHG's row insertion, scans, compaction and
// elimination into the matrix's PHG-order arrays (assemble.hip refill plan),
// the element operator's coefficients and the diagonal in the solver
// numbering -- then the start of the reliable-update solver and CH of its
// slots; the second graph holds CH further slots.  PHG drops exact zeros, so
// the pattern may change with the coefficients: nothing in the graphs
// depends on nnz.  The FP32 iterations run on the element operator (its
// layout is a function of the mesh alone) and the FP64 residuals on PHG's
// CSR through the solver numbering (PhgPermView).  The slots read their
// system's state on the device and return at once once it is done: the
// host launches the first graph and then the second, two launches in
// flight, and reads the states each launch copies to pinned memory only
// after that launch has ended, while the next one runs.
struct phg_hip_batch {
    struct Item {
        phg_hip_mat *A = nullptr;       // the caller's matrix, re-assembled in place
        phg_hip_space *sp = nullptr;
        phg_hip_problem pb;
        const double *kappa = nullptr;
        phg::RefillPlan *plan = nullptr;
        // the solver's view of A: A's PHG-order arrays (not owned) with the
        // batch's solver numbering and its inverse, diagonals in that
        // numbering, a workspace, and A's element operator (handed back to
        // A when the batch is destroyed)
        phg_hip_mat S;
        phg_hip_vec b;                  // view of A's right-hand side
        phg_hip_vec *x = nullptr;
        phg_hip_vec *inputs[3] = {};    // kappa (or null), f, g
        MultiItem run;
        bool held = false;              // A, sp and the vectors count this batch
    };
    std::vector<Item> items;
    hipGraphExec_t first = nullptr, next = nullptr;
    void *pinned = nullptr;             // [nmat] State, then [nmat] int64 entry counts
    int32_t maxit = 0;
    // the matrix, space and vectors of an item count the batch (+1 / -1)
    static void hold(Item &it, int d)
    {
        it.A->batch_refs += d;
        it.sp->batch_refs += d;
        it.x->batch_refs += d;
        for (phg_hip_vec *v : it.inputs)
            if (v)
                v->batch_refs += d;
    }
    ~phg_hip_batch()
    {
        if (first)
            (void)hipGraphExecDestroy(first);
        if (next)
            (void)hipGraphExecDestroy(next);
        if (pinned)
            (void)hipHostFree(pinned);
        for (Item &it : items) {
            if (it.held)
                hold(it, -1);
            phg::refill_plan_destroy(it.plan);
            phg_hip_mat &S = it.S;
            phg::free_workspace(&S);
            // the operator's coefficients are those of the last assembly, as
            // are A's values; its numbering is A's solver numbering, which a
            // rebuild of A's solver copy reproduces
            if (S.eop != nullptr && it.A != nullptr && it.A->eop == nullptr)
                it.A->eop = S.eop;
            else if (S.eop != nullptr)
                phg::free_eop(&S);
            S.eop = nullptr;
            phg::dfree(S.q_perm);
            phg::dfree(S.q_iperm);
            phg::dfree(S.q_diag);
            phg::dfree(S.q_diag32);
            S.q_perm = S.q_iperm = nullptr;
            S.q_diag = nullptr;
            S.q_diag32 = nullptr;
        }
    }
};

namespace {

constexpr int BATCH_CH = 16;   // slots per graph launch (even: the state parity holds)

__global__ void k_perm_init(int64_t n, const int32_t *perm, int32_t *q, int32_t *iq)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        const int32_t r = perm ? perm[i] : (int32_t)i;
        q[i] = r;
        iq[r] = (int32_t)i;
    }
}

// the diagonal in the solver numbering, FP64 and FP32
__global__ void k_diag_gather(int64_t n, const int32_t *perm, const double *d, double *q,
                              float *q32)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        const double v = d[perm[i]];
        q[i] = v;
        q32[i] = (float)v;
    }
}

// captures one graph: per system on a pool stream, (assembly + solver start
// when first) and BATCH_CH slots, then the copy of the state (and, when
// first, of the entry count) to pinned memory
int batch_capture(phg_hip_batch *B, bool first, double rtol, double btol, double atol,
                  double delta, hipGraphExec_t *out)
{
    hipStream_t s = phg::ctx()->stream;
    hipStream_t *pool = nullptr;
    int npool = 0, rc;
    if ((rc = phg::aux_streams(&pool, &npool)))
        return rc;
    const int nmat = (int)B->items.size();
    State *hm = (State *)B->pinned;
    int64_t *hn = (int64_t *)(hm + nmat);
    hipEvent_t fork = nullptr, join[MULTI_MAX] = {};
    hipGraph_t graph = nullptr;
    auto cleanup = [&]() {
        if (fork)
            (void)hipEventDestroy(fork);
        for (int i = 0; i < nmat; i++)
            if (join[i])
                (void)hipEventDestroy(join[i]);
        if (graph)
            (void)hipGraphDestroy(graph);
    };
    hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    for (int i = 0; i < nmat && e == hipSuccess; i++)
        e = hipEventCreateWithFlags(&join[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        cleanup();
        return phg::hip_fail(e, "batch events", __FILE__, __LINE__);
    }
    if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed)) != hipSuccess) {
        cleanup();
        return phg::hip_fail(e, "batch capture", __FILE__, __LINE__);
    }
    rc = PHG_HIP_OK;
    e = hipEventRecord(fork, s);
    for (int i = 0; i < nmat && e == hipSuccess && rc == PHG_HIP_OK; i++) {
        phg_hip_batch::Item &it = B->items[(size_t)i];
        phg_hip_mat &S = it.S;
        hipStream_t si = pool[i % npool];
        if ((e = hipStreamWaitEvent(si, fork, 0)) != hipSuccess)
            break;
        if (first) {
            if ((rc = phg::refill_enqueue(it.plan, si)))
                break;
            hipLaunchKernelGGL(k_diag_gather, dim3(grid_for(S.n)), dim3(BS), 0, si, S.n,
                               (const int32_t *)S.q_perm, (const double *)it.A->d_diag,
                               S.q_diag, S.q_diag32);
            if ((rc = phg::eop_coefficients(it.sp, it.pb.use_kappa ? it.kappa : nullptr,
                                            S.eop->lay->eperm, it.pb.mass_coef, S.eop->coef,
                                            si)))
                break;
            auto pl = std::make_shared<RuPlan<PhgPermView>>();
            if ((rc = ru_prologue<PhgPermView>(&S, &it.b, it.x, B->maxit, rtol, btol, atol,
                                               delta, si, pl.get(), true)))
                break;
            phg_hip_mat *Sp = &S;
            phg_hip_vec *x = it.x;
            it.run.slot = [Sp, pl](int64_t k, hipStream_t st) {
                ru_slot<PhgPermView>(Sp, *pl, k, st, false);
            };
            it.run.finish = [Sp, pl, x](hipStream_t st) { ru_finish<PhgPermView>(Sp, *pl, x, st); };
            it.run.st = pl->ws->st;
        }
        for (int k = 0; k < BATCH_CH; k++)
            it.run.slot(k, si);
        if ((e = hipMemcpyAsync(hm + i, it.run.st + (BATCH_CH & 1), sizeof(State),
                                hipMemcpyDeviceToHost, si)) != hipSuccess)
            break;
        if (first && (e = hipMemcpyAsync(hn + i, phg::refill_nnz(it.plan), sizeof(int64_t),
                                         hipMemcpyDeviceToHost, si)) != hipSuccess)
            break;
        if ((e = hipEventRecord(join[i], si)) != hipSuccess ||
            (e = hipStreamWaitEvent(s, join[i], 0)) != hipSuccess)
            break;
    }
    if (e == hipSuccess)
        e = hipGetLastError();
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (rc == PHG_HIP_OK && e == hipSuccess && e2 == hipSuccess)
        e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    cleanup();
    if (rc)
        return rc;
    if (e != hipSuccess || e2 != hipSuccess)
        return phg::hip_fail(e != hipSuccess ? e : e2, "batch graph capture", __FILE__,
                             __LINE__);
    return PHG_HIP_OK;
}

// the solver's view of item it's matrix (see phg_hip_batch::Item)
int batch_item_setup(phg_hip_batch::Item &it, const phg_hip_vec *f, const phg_hip_vec *g)
{
    phg_hip_mat *A = it.A;
    phg_hip_mat &S = it.S;
    hipStream_t s = phg::ctx()->stream;
    const int64_t n = A->n;
    int rc;
    if ((rc = phg::ensure_solver_copy(A)) ||
        (rc = phg::dmalloc((void **)&S.q_perm, (size_t)n * 4 + 4)) ||
        (rc = phg::dmalloc((void **)&S.q_iperm, (size_t)n * 4 + 4)) ||
        (rc = phg::dmalloc((void **)&S.q_diag, (size_t)n * 8 + 8)) ||
        (rc = phg::dmalloc((void **)&S.q_diag32, (size_t)n * 4 + 4)))
        return rc;
    hipLaunchKernelGGL(k_perm_init, dim3(grid_for(n)), dim3(BS), 0, s, n,
                       (const int32_t *)A->q_perm, S.q_perm, S.q_iperm);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    S.eop = A->eop;
    A->eop = nullptr;
    S.eop->on = 1;
    // the plan re-allocates A's columns and values and drops its solver copy
    if ((rc = phg::refill_plan_create(it.sp, &it.pb, A, it.kappa, f->d, g->d, &it.plan)))
        return rc;
    S.n = n;
    S.dim = A->dim;
    S.ndof = A->ndof;
    S.nnz = A->nnz;
    S.eliminated = 1;
    S.d_rowptr = A->d_rowptr;
    S.d_cols = A->d_cols;
    S.d_vals = A->d_vals;
    S.d_rhs = A->d_rhs;
    S.d_diag = A->d_diag;
    if (get_ws(&S, 1) == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "workspace allocation failed");
    it.b.n = n;
    it.b.d = A->d_rhs;
    return PHG_HIP_OK;
}

}  // namespace

extern "C" int phg_hip_batch_create(int32_t nmat, phg_hip_space *const *sp,
                                    const phg_hip_problem *const *pb,
                                    phg_hip_mat *const *Av, const phg_hip_vec *const *kappa,
                                    const phg_hip_vec *const *f, const phg_hip_vec *const *g,
                                    phg_hip_vec *const *x, int32_t maxit, double rtol,
                                    double btol, double atol, double delta,
                                    phg_hip_batch **out)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (out == nullptr || nmat < 1 || nmat > MULTI_MAX || sp == nullptr || pb == nullptr ||
        Av == nullptr || f == nullptr || g == nullptr || x == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_batch_create: 1 <= nmat <= %d and "
                                             "non-null arrays required", MULTI_MAX);
    if (!(delta > 0.0 && delta < 1.0))
        return phg::set_error(PHG_HIP_E_ARG, "delta must lie in (0, 1)");
    *out = nullptr;
    for (int i = 0; i < nmat; i++) {
        phg_hip_mat *A = Av[i];
        if (sp[i] == nullptr || pb[i] == nullptr || A == nullptr || f[i] == nullptr ||
            g[i] == nullptr || x[i] == nullptr)
            return phg::set_error(PHG_HIP_E_ARG, "phg_hip_batch_create: null entry %d", i);
        for (int j = 0; j < i; j++)
            if (Av[j] == A)
                return phg::set_error(PHG_HIP_E_ARG,
                                      "phg_hip_batch_create: a matrix appears twice");
        const phg_hip_problem &p = *pb[i];
        const int64_t ne = sp[i]->nelem, nf = sp[i]->ndof * p.dim;
        if ((p.use_kappa && (kappa == nullptr || kappa[i] == nullptr ||
                             kappa[i]->n != ne)) ||
            f[i]->n != nf || g[i]->n != A->n || x[i]->n != A->n)
            return phg::set_error(PHG_HIP_E_ARG,
                                  "phg_hip_batch_create: system %d: kappa needs nelem, f "
                                  "ndof*dim, g and x n entries", i);
        if (A->eop == nullptr || A->eop->nelem != ne ||
            (p.mass_coef != 0.0) != (A->eop->mass != 0) || A->space_uid != sp[i]->uid)
            return phg::set_error(PHG_HIP_E_ARG,
                                  "phg_hip_batch_create: system %d needs the eliminated "
                                  "matrix assembled on its space with its element operator "
                                  "(phg_hip_mat_element_operator) for this problem", i);
    }
    std::unique_ptr<phg_hip_batch> B(new phg_hip_batch());
    B->items.resize((size_t)nmat);
    B->maxit = maxit;
    for (int i = 0; i < nmat; i++) {
        phg_hip_batch::Item &it = B->items[(size_t)i];
        const phg_hip_problem &p = *pb[i];
        it.A = Av[i];
        it.sp = sp[i];
        it.pb = p;
        it.kappa = p.use_kappa ? kappa[i]->d : nullptr;
        it.x = x[i];
        it.inputs[0] = p.use_kappa ? const_cast<phg_hip_vec *>(kappa[i]) : nullptr;
        it.inputs[1] = const_cast<phg_hip_vec *>(f[i]);
        it.inputs[2] = const_cast<phg_hip_vec *>(g[i]);
        if ((rc = batch_item_setup(it, f[i], g[i])))
            return rc;
    }
    if (hipHostMalloc(&B->pinned, (size_t)nmat * (sizeof(State) + sizeof(int64_t)) + 64,
                      hipHostMallocDefault) != hipSuccess)
        return phg::set_error(PHG_HIP_E_NOMEM, "pinned mailbox allocation failed");
    PHG_HIP_CHECK(hipStreamSynchronize(phg::ctx()->stream));
    if ((rc = batch_capture(B.get(), true, rtol, btol, atol, delta, &B->first)) ||
        (rc = batch_capture(B.get(), false, rtol, btol, atol, delta, &B->next)))
        return rc;
    for (auto &it : B->items) {
        phg_hip_batch::hold(it, +1);
        it.held = true;
    }
    *out = B.release();
    return PHG_HIP_OK;
}

extern "C" int phg_hip_batch_run(phg_hip_batch *B, int32_t *nits, double *residual)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (B == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null batch");
    const int nmat = (int)B->items.size();
    hipStream_t s = phg::ctx()->stream;
    State *hm = (State *)B->pinned;
    const int64_t *hn = (const int64_t *)(hm + nmat);
    constexpr int D = 2;                 // graph launches in flight
    hipEvent_t ev[D] = {};
    auto cleanup = [&]() {
        for (int k = 0; k < D; k++)
            if (ev[k])
                (void)hipEventDestroy(ev[k]);
    };
    hipError_t e = hipSuccess;
    for (int k = 0; k < D && e == hipSuccess; k++)
        e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
    int64_t launched = 0, ended = 0;
    auto launch = [&]() -> hipError_t {
        hipError_t e1 = hipGraphLaunch(launched == 0 ? B->first : B->next, s);
        if (e1 == hipSuccess)
            e1 = hipEventRecord(ev[launched % D], s);
        launched++;
        return e1;
    };
    const int64_t limit = (2 * (int64_t)B->maxit + 4096) / BATCH_CH + D;
    bool stuck = false;
    while (e == hipSuccess && launched < D)
        e = launch();
    while (e == hipSuccess) {
        if ((e = hipEventSynchronize(ev[ended % D])) != hipSuccess)
            break;
        ended++;
        bool all = true;
        for (int i = 0; i < nmat; i++)
            all = all && hm[i].done != RUNNING;
        if (all)
            break;
        if (launched > limit) {
            stuck = true;
            break;
        }
        e = launch();
    }
    if (e == hipSuccess && !stuck) {
        for (auto &it : B->items)
            it.run.finish(s);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess)
        return phg::hip_fail(e, "batch run", __FILE__, __LINE__);
    // the matrices: new entry counts; the row blocks of the PHG-order CSR and
    // any solver copy built since the last run follow the new pattern when
    // next needed
    for (int i = 0; i < nmat; i++) {
        phg_hip_mat *A = B->items[(size_t)i].A;
        A->nnz = hn[i];
        B->items[(size_t)i].S.nnz = hn[i];
        phg::dfree(A->d_rblk);
        A->d_rblk = nullptr;
        A->nrblk = 0;
        phg::free_solver_copy(A);
        phg::free_sell(A);
        A->q_stale = 1;
    }
    if (stuck)
        return phg::set_error(PHG_HIP_E_HIP, "batched solve did not finish");
    int32_t mx = 0;
    for (int i = 0; i < nmat; i++) {
        if (hm[i].done == ZERO_DIAG)
            return phg::set_error(PHG_HIP_E_ZERO_DIAG, "zero diagonal in Jacobi preconditioner");
        if (nits)
            nits[i] = hm[i].nits;
        if (residual)
            residual[i] = hm[i].res;
        mx = std::max(mx, hm[i].nits);
    }
    return mx;
}

extern "C" int phg_hip_batch_destroy(phg_hip_batch *B)
{
    if (B == nullptr)
        return PHG_HIP_OK;
    if (phg::ctx())
        (void)hipStreamSynchronize(phg::ctx()->stream);
    delete B;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_mat_set_format(phg_hip_mat *A, int32_t fmt64,
                                      int32_t fmt32)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr || !A->eliminated || fmt64 < 0 || fmt64 >= FMT_N ||
        fmt32 < 0 || fmt32 >= FMT_N)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_mat_set_format: bad args");
    A->fmt64 = fmt64;
    A->fmt32 = fmt32;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_mat_get_format(const phg_hip_mat *A, int32_t *fmt64,
                                      int32_t *fmt32)
{
    if (A == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null matrix");
    if (fmt64)
        *fmt64 = A->fmt64;
    if (fmt32)
        *fmt32 = A->fmt32;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_element_apply(const phg_hip_mat *A_, const phg_hip_vec *x,
                                     phg_hip_vec *y)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    if (A == nullptr || A->eop == nullptr || x == nullptr || y == nullptr ||
        x->n != A->n || y->n != A->n || x->d == y->d)
        return phg::set_error(PHG_HIP_E_ARG,
                              "phg_hip_element_apply: operator not built, size "
                              "mismatch or x aliases y");
    // the operator's Dirichlet rows read the FP32 diagonal
    if ((rc = phg::ensure_solver_copy(A)) || (rc = phg::ensure_fp32_copy(A)))
        return rc;
    hipStream_t s = phg::ctx()->stream;
    Workspace *ws = get_ws(A, 1);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "workspace allocation failed");
    const int64_t n = A->n;
    const int g = grid_for(n);
    hipLaunchKernelGGL(k_gather_f32, dim3(g), dim3(BS), 0, s, n,
                       (const int32_t *)A->q_perm, (const double *)x->d, ws->p32);
    phg::eop_apply(A->eop, ws->p32, (const float *)A->q_diag32, ws->Ap32, ws->seg(SEG_S),
                   nullptr, nullptr, 0, s);
    hipLaunchKernelGGL(k_scatter_f32, dim3(g), dim3(BS), 0, s, n,
                       (const int32_t *)A->q_perm, (const float *)ws->Ap32, y->d);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    return PHG_HIP_OK;
}

extern "C" int phg_hip_element_time(const phg_hip_mat *A_, int32_t reps,
                                    double *elem_us, double *gather_us)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    phg_hip_mat *A = const_cast<phg_hip_mat *>(A_);
    if (A == nullptr || A->eop == nullptr || elem_us == nullptr || gather_us == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_element_time: operator not built");
    if ((rc = phg::ensure_solver_copy(A)) || (rc = phg::ensure_fp32_copy(A)))
        return rc;
    if (reps <= 0)
        reps = 10;
    hipStream_t s = phg::ctx()->stream;
    Workspace *ws = get_ws(A, 1);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "workspace allocation failed");
    PHG_HIP_CHECK(hipMemsetAsync(ws->p32, 0, (size_t)A->n * sizeof(float), s));
    hipEvent_t ev[3];
    for (int i = 0; i < 3; i++)
        PHG_HIP_CHECK(hipEventCreate(&ev[i]));
    double t[2] = {0.0, 0.0};
    hipError_t err = hipSuccess;
    for (int k = 0; k < 2 && err == hipSuccess; k++) {
        for (int i = 0; i < reps + 2; i++) {
            if (i == 2)
                err = hipEventRecord(ev[0], s);
            phg::eop_apply(A->eop, ws->p32, (const float *)A->q_diag32, ws->Ap32,
                           ws->seg(SEG_S), nullptr, nullptr, 0, s, k + 1);
        }
        if (err == hipSuccess)
            err = hipGetLastError();
        if (err == hipSuccess)
            err = hipEventRecord(ev[1], s);
        if (err == hipSuccess)
            err = hipEventSynchronize(ev[1]);
        float ms = 0.f;
        if (err == hipSuccess)
            err = hipEventElapsedTime(&ms, ev[0], ev[1]);
        t[k] = 1e3 * ms / reps;
    }
    for (int i = 0; i < 3; i++)
        (void)hipEventDestroy(ev[i]);
    if (err != hipSuccess)
        return phg::set_error(PHG_HIP_E_HIP, "element operator timing: %s",
                              hipGetErrorString(err));
    *elem_us = t[0];
    *gather_us = t[1];
    return PHG_HIP_OK;
}

extern "C" int phg_hip_element_bytes(const phg_hip_mat *A, int64_t *elem_bytes,
                                     int64_t *gather_bytes)
{
    if (A == nullptr || A->eop == nullptr || elem_bytes == nullptr ||
        gather_bytes == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_element_bytes: operator not built");
    const EopData *eo = A->eop;
    const int64_t nsh = eo->nent - eo->nint;    // shared-row entries = partials
    // block kernel: packed indices and coefficients of every element, per
    // touched-row entry its row (and partial slot when shared), slot start
    // and x; interior rows of y and the partials written; block tables;
    // Dirichlet rows (row, x, diagonal read, y written)
    *elem_bytes = eo->nelem * (4 * phg::eop_record_dwords(eo->p) + 4 * (eo->mass ? 7 : 6)) +
                  eo->nent * (4 + 2 + 4) +
                  eo->nblk * (2 + 3 * 4) + nsh * (4 + 4) + eo->nint * 4 + eo->ndir * 16;
    // fix-up: rows, partial ranges and partials of the shared rows, y
    *gather_bytes = eo->nfx * (4 + 4 + 4) + 4 + eo->npart * 4;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_element_layout(const phg_hip_mat *A, int64_t out[7])
{
    if (A == nullptr || A->eop == nullptr || out == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_element_layout: operator not built");
    const EopData *eo = A->eop;
    out[0] = eo->nblk;
    out[1] = eo->nent;
    out[2] = eo->nint;
    out[3] = eo->nfx;
    out[4] = eo->npart;
    out[5] = eo->ndir;
    out[6] = eo->lay && eo->lay->on_device ? 1 : 0;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_mat_tune(phg_hip_mat *A, int32_t reps)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr || !A->eliminated)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_mat_tune: matrix not assembled");
    if (reps <= 0)
        reps = 10;
    if ((rc = ensure_format(A, FMT_SELL, 1)) || (rc = phg::build_dict(A)) ||
        (rc = phg::build_xstage(A)))
        return rc;
    for (int f = 0; f < FMT_N; f++) {
        double t64 = 0, t32 = 0;
        if (((f == FMT_DICT || f == FMT_DICT_NP) && A->cd.state != 1) ||
            ((f == FMT_XS || f == FMT_XS_PF) && A->xs.state != 1)) {
            A->tune_us[f][0] = A->tune_us[f][1] = -1.0;   // not representable
            continue;
        }
        if ((rc = ensure_format(A, f, 1)))
            return rc;
        rc = with_format(f, [&](auto tag) {
            using View = decltype(tag);
            using F = FmtData<View>;
            int e;
            if ((e = time_spmv<double, View>(A, F::v64(A), reps, &t64)) ||
                (e = time_spmv<float, View>(A, F::v32(A), reps, &t32)))
                return e;
            return (int)PHG_HIP_OK;
        });
        if (rc)
            return rc;
        A->tune_us[f][0] = t64;
        A->tune_us[f][1] = t32;
    }
    int b64 = 0, b32 = 0;
    for (int f = 1; f < FMT_N; f++) {
        if (A->tune_us[f][0] < 0.0)
            continue;
        if (A->tune_us[f][0] < A->tune_us[b64][0])
            b64 = f;
        if (A->tune_us[f][1] < A->tune_us[b32][1])
            b32 = f;
    }
    A->fmt64 = b64;
    A->fmt32 = b32;
    // copies no solver uses are released
    if (b64 != FMT_SELL && b32 != FMT_SELL)
        phg::free_sell(A);
    auto uses = [&](int a, int b) { return b64 == a || b32 == a || b64 == b || b32 == b; };
    if (!uses(FMT_DICT, FMT_DICT_NP))
        phg::free_dict_layout(&A->cd);
    if (!uses(FMT_XS, FMT_XS_PF))
        phg::free_dict_layout(&A->xs);
    return PHG_HIP_OK;
}

extern "C" int phg_hip_mat_format_bytes(const phg_hip_mat *A, int32_t fmt,
                                        int32_t fp32, int64_t *bytes)
{
    if (A == nullptr || fmt < 0 || fmt >= FMT_N || bytes == nullptr ||
        !A->eliminated)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_mat_format_bytes: bad args");
    const int64_t sv = fp32 ? 4 : 8, n = A->n;
    const int64_t vec = 2 * n * sv;            // one read of x, one write of y
    switch (fmt) {
    case FMT_SELL:
        if (A->d_svals == nullptr)
            return phg::set_error(PHG_HIP_E_STATE, "SELL layout not built");
        *bytes = A->sell_nnz * (sv + 4) + A->nslice * 12 + vec;
        break;
    case FMT_DICT:
    case FMT_DICT_NP:
    case FMT_XS:
    case FMT_XS_PF: {
        const DictLayout &L = (fmt == FMT_XS || fmt == FMT_XS_PF) ? A->xs : A->cd;
        if (L.state != 1)
            return phg::set_error(PHG_HIP_E_STATE, "dictionary layout not built");
        *bytes = L.nent * (sv + 2) + (n + 1) * 4 + (L.nblk + 1) * 12 +
                 L.ndict * 4 + vec;
        break;
    }
    default:
        *bytes = A->nnz * (sv + 4) + (n + 1) * 4 + vec;
    }
    return PHG_HIP_OK;
}

extern "C" int phg_hip_mat_tune_time(const phg_hip_mat *A, int32_t fmt,
                                     int32_t fp32, double *us)
{
    if (A == nullptr || fmt < 0 || fmt >= FMT_N || us == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_mat_tune_time: bad args");
    *us = A->tune_us[fmt][fp32 ? 1 : 0];
    return PHG_HIP_OK;
}

extern "C" int phg_hip_stream_copy(int64_t n, int32_t reps, double *us)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (n <= 0 || reps <= 0 || us == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_stream_copy: bad args");
    hipStream_t s = phg::ctx()->stream;
    double *a = nullptr, *b = nullptr;
    if ((rc = phg::dmalloc((void **)&a, (size_t)n * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&b, (size_t)n * sizeof(double)))) {
        phg::dfree(a);
        return rc;
    }
    PHG_HIP_CHECK(hipMemsetAsync(a, 0, (size_t)n * sizeof(double), s));
    int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    int g = (int)std::min<int64_t>((n + BS - 1) / BS, (int64_t)cus * 16);
    hipEvent_t e0, e1;
    PHG_HIP_CHECK(hipEventCreate(&e0));
    PHG_HIP_CHECK(hipEventCreate(&e1));
    for (int i = 0; i < reps + 1; i++) {
        if (i == 1)
            PHG_HIP_CHECK(hipEventRecord(e0, s));
        hipLaunchKernelGGL(k_copy, dim3(g), dim3(BS), 0, s, n,
                           (const double *)a, b);
    }
    hipError_t err = hipGetLastError();
    float ms = 0.f;
    if (err == hipSuccess)
        err = hipEventRecord(e1, s);
    if (err == hipSuccess)
        err = hipEventSynchronize(e1);
    if (err == hipSuccess)
        err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    phg::dfree(a);
    phg::dfree(b);
    if (err != hipSuccess)
        return phg::set_error(PHG_HIP_E_HIP, "copy timing: %s", hipGetErrorString(err));
    *us = 1e3 * ms / reps;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_profile(int32_t on)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    PHG_HIP_CHECK(hipStreamSynchronize(phg::ctx()->stream));
    prof_collect();
    for (int k = 0; k < KP_N; k++) {
        g_prof.launches[k] = 0;
        g_prof.ms[k] = 0.0;
    }
    g_prof.on = on != 0;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_profile_read(int32_t kind, int64_t *launches, double *ms)
{
    if (kind < 0 || kind >= KP_N)
        return phg::set_error(PHG_HIP_E_ARG, "profile kind %d out of range", kind);
    if (launches)
        *launches = g_prof.launches[kind];
    if (ms)
        *ms = g_prof.ms[kind];
    return PHG_HIP_OK;
}
