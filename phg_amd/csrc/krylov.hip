// Krylov kernels for gfx950: CSR SpMV, BLAS-1 and the Jacobi-PCG of
// solver-phg.c:485 restructured into three fused kernels per iteration.
//
// Per iteration (vector streams in brackets):
//   k_pcg_p      p = z + beta*p with z = r/diag computed on the fly    [r,d,p,p]
//   k_pcg_spmv   Ap = A p and the partial sums of (p, Ap)              [A,p,Ap]
//   k_pcg_update x += alpha p, r -= alpha Ap, partial sums of (r,r) and
//                (r, r/diag) for the next iteration                    [x,p,r,Ap,d,x,r]
// Scalars (alpha, beta, residual, iteration count) never leave the device:
// the last work-group to finish a kernel reduces the per-group partial sums
// in a fixed order (deterministic) and updates the solver state, and every
// kernel returns at once when the state says the solve is over.  The host
// launches iterations in chunks and reads the state from a pinned mailbox
// once per chunk.
//
// Work-groups are remapped so that each of the 8 XCDs (own L2) processes a
// contiguous slice of rows, keeping the gathered x entries of neighbouring
// rows in the same L2.
#include "phg_hip_internal.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>

namespace {

constexpr int BS = 256;      // work-group size (4 wavefronts of 64)
constexpr int NXCD = 8;

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

// Block reduction of NV values; the partials of every work-group go to
// partials[k * gridDim.x + block]; the last work-group to arrive sums them
// in a fixed order and returns true with the totals in out[] (all threads).
template <int NV>
__device__ bool reduce_last(double (&v)[NV], double *partials,
                            unsigned *counter, double (&out)[NV])
{
    __shared__ double sh[NV][BS / 64];
    __shared__ int last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0)
            sh[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < BS / 64; w++)
                s += sh[k][w];
            partials[(int64_t)k * gridDim.x + blockIdx.x] = s;
        }
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last)
        return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++)
        acc[k] = 0.0;
    for (int64_t i = threadIdx.x; i < gridDim.x; i += BS)
#pragma unroll
        for (int k = 0; k < NV; k++)
            acc[k] += __hip_atomic_load(&partials[(int64_t)k * gridDim.x + i],
                                        __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
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
    if (threadIdx.x == 0)
        *counter = 0u;
    return true;
}

// solver state kept in device memory
struct State {
    double rz, rz_prev, pAp, alpha;
    double res, tol, rb, r0;
    double rtol, btol, atol, scale;
    int32_t nits, done, maxit, xnz;
    int32_t zero_diag, outer, pad0, pad1;
    unsigned counter[4];
};
// done codes
constexpr int RUNNING = 0, CONVERGED = 1, BREAKDOWN = 2, ZERO_DIAG = 3;

// IsZero (utils.h:451)
__device__ inline bool is_zero(double a) { return fabs(a) < 10.0 * 2.2250738585072014e-308; }

// contiguous row range of this work-group
__device__ inline void row_range(int64_t n, int64_t &r0, int64_t &r1)
{
    int64_t b = xcd_block(), nb = gridDim.x;
    int64_t per = (n + nb - 1) / nb;
    r0 = b * per;
    r1 = r0 + per < n ? r0 + per : n;
}

// sum over a row with L lanes (result valid in every lane of the group)
template <int L, typename VT, typename XT, typename AT>
__device__ inline AT row_dot(const int32_t *__restrict__ rp,
                             const int32_t *__restrict__ ci,
                             const VT *__restrict__ va,
                             const XT *__restrict__ x, int64_t row, int lane)
{
    AT s = 0;
    int32_t j0 = rp[row], j1 = rp[row + 1];
    for (int32_t j = j0 + lane; j < j1; j += L)
        s += (AT)va[j] * (AT)x[ci[j]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1)
        s += __shfl_xor(s, o, L);
    return s;
}

//------------------------------------------------------------- plain SpMV
template <int L>
__global__ void __launch_bounds__(BS) k_spmv64(int64_t n, const int32_t *rp,
                                               const int32_t *ci,
                                               const double *va,
                                               const double *x, double *y,
                                               double alpha, double beta)
{
    int64_t r0, r1;
    row_range(n, r0, r1);
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    for (int64_t row = r0 + grp; row < r1; row += BS / L) {
        double s = row_dot<L, double, double, double>(rp, ci, va, x, row, lane);
        if (lane == 0) {
            // matvec.c:608-633 beta first, then alpha * A x
            double yo = beta == 0.0 ? 0.0 : beta * y[row];
            y[row] = yo + alpha * s;
        }
    }
}

//------------------------------------------------------------------ BLAS-1
__global__ void __launch_bounds__(BS) k_dot(int64_t n, const double *x,
                                            const double *y, double *partials,
                                            unsigned *counter, double *out)
{
    double v[1] = {0.0}, o[1];
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        v[0] += x[i] * y[i];
    if (reduce_last<1>(v, partials, counter, o) && threadIdx.x == 0)
        *out = o[0];
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
// Prologue (solver-phg.c:505-533): |b|, initial residual, tolerance, and
// (r, z) for the first iteration.  r = b - A x when |x|_1/n is not zero.
template <int L>
__global__ void __launch_bounds__(BS) k_pcg_init_norms(int64_t n,
                                                       const double *b,
                                                       const double *x,
                                                       double *partials,
                                                       State *st)
{
    double v[2] = {0.0, 0.0}, o[2];
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        v[0] += b[i] * b[i];
        v[1] += fabs(x[i]);
    }
    if (reduce_last<2>(v, partials, &st->counter[0], o) && threadIdx.x == 0) {
        st->rb = sqrt(o[0]);
        st->xnz = !is_zero(o[1] / (double)n);
    }
}

template <int L>
__global__ void __launch_bounds__(BS) k_pcg_init_res(
    int64_t n, const int32_t *rp, const int32_t *ci, const double *va,
    const double *diag, const double *b, double *x, double *r,
    double *partials, State *st)
{
    int64_t r0, r1;
    row_range(n, r0, r1);
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    const bool bzero = is_zero(st->rb);
    const bool xnz = st->xnz != 0;
    double v[3] = {0.0, 0.0, 0.0}, o[3];
    for (int64_t row = r0 + grp; row < r1; row += BS / L) {
        double s = 0.0;
        if (xnz && !bzero)
            s = row_dot<L, double, double, double>(rp, ci, va, x, row, lane);
        if (lane == 0) {
            double ri;
            if (bzero) {
                x[row] = 0.0;
                ri = 0.0;
            }
            else
                ri = xnz ? -s + b[row] : b[row];
            r[row] = ri;
            double d = diag[row];
            v[0] += ri * ri;
            if (d == 0.0)
                v[2] += 1.0;
            else
                v[1] += ri * (ri / d);
        }
    }
    if (reduce_last<3>(v, partials, &st->counter[0], o) && threadIdx.x == 0) {
        st->nits = 0;
        if (bzero) {
            st->res = 0.0;
            st->done = CONVERGED;
            return;
        }
        double res = sqrt(o[0]);
        st->r0 = st->res = res;
        double tol = res * st->rtol;   // get_tol (solver-phg.c:104)
        if (tol < st->rb * st->btol)
            tol = st->rb * st->btol;
        if (tol < st->atol)
            tol = st->atol;
        st->tol = tol;
        st->rz = o[1];
        st->rz_prev = 0.0;
        bool run = res > tol && st->maxit > 0;
        st->zero_diag = o[2] != 0.0;
        st->done = !run ? CONVERGED : (o[2] != 0.0 ? ZERO_DIAG : RUNNING);
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_pcg_p(int64_t n, const T *r,
                                              const T *diag, T *p,
                                              const State *st)
{
    if (st->done)
        return;
    const bool first = st->nits == 0;
    const T beta = first ? (T)0 : (T)(st->rz / st->rz_prev);
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        T z = r[i] / diag[i];
        p[i] = first ? z : beta * p[i] + z;
    }
}

template <int L, typename T>
__global__ void __launch_bounds__(BS) k_pcg_spmv(
    int64_t n, const int32_t *rp, const int32_t *ci, const T *va, const T *p,
    T *Ap, double *partials, State *st)
{
    if (st->done)
        return;
    int64_t r0, r1;
    row_range(n, r0, r1);
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    double v[1] = {0.0}, o[1];
    for (int64_t row = r0 + grp; row < r1; row += BS / L) {
        T s = row_dot<L, T, T, T>(rp, ci, va, p, row, lane);
        if (lane == 0) {
            Ap[row] = s;
            v[0] += (double)p[row] * (double)s;
        }
    }
    if (reduce_last<1>(v, partials, &st->counter[1], o) && threadIdx.x == 0) {
        st->nits += 1;
        st->pAp = o[0];
        if (o[0] == 0.0)
            st->done = BREAKDOWN;      // solver-phg.c:561
        else
            st->alpha = st->rz / o[0];
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_pcg_update(int64_t n, T *x,
                                                   const T *p, T *r,
                                                   const T *Ap, const T *diag,
                                                   double *partials,
                                                   State *st)
{
    if (st->done)
        return;
    const T alpha = (T)st->alpha;
    double v[2] = {0.0, 0.0}, o[2];
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        x[i] += alpha * p[i];
        T ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        T z = ri / diag[i];
        v[0] += (double)ri * (double)ri;
        v[1] += (double)ri * (double)z;
    }
    if (reduce_last<2>(v, partials, &st->counter[2], o) && threadIdx.x == 0) {
        double res = sqrt(o[0]);
        st->res = res;
        st->rz_prev = st->rz;
        st->rz = o[1];
        if (!(res > st->tol && st->nits < st->maxit))
            st->done = CONVERGED;
        else if (st->rz_prev == 0.0) {  // solver-phg.c:547, next iteration
            st->nits += 1;
            st->done = BREAKDOWN;
        }
    }
}

//----------------------------------------------- mixed-precision refinement
// r = b - A x in FP64 with |r|^2 and |b|^2; sets the outer state.
template <int L>
__global__ void __launch_bounds__(BS) k_ir_residual(
    int64_t n, const int32_t *rp, const int32_t *ci, const double *va,
    const double *b, const double *x, double *r, double *partials, State *so)
{
    int64_t r0, r1;
    row_range(n, r0, r1);
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    double v[2] = {0.0, 0.0}, o[2];
    for (int64_t row = r0 + grp; row < r1; row += BS / L) {
        double s = row_dot<L, double, double, double>(rp, ci, va, x, row, lane);
        if (lane == 0) {
            double bi = b[row], ri = bi - s;
            r[row] = ri;
            v[0] += ri * ri;
            v[1] += bi * bi;
        }
    }
    if (reduce_last<2>(v, partials, &so->counter[0], o) && threadIdx.x == 0) {
        double res = sqrt(o[0]);
        so->res = res;
        if (so->outer == 0) {
            so->rb = sqrt(o[1]);
            so->r0 = res;
            double tol = res * so->rtol;
            if (tol < so->rb * so->btol)
                tol = so->rb * so->btol;
            if (tol < so->atol)
                tol = so->atol;
            so->tol = tol;
        }
        so->done = (res <= so->tol || is_zero(so->rb)) ? CONVERGED : RUNNING;
    }
}

// inner right-hand side r32 = r / |r|, x32 = 0, inner state for the FP32 PCG
__global__ void __launch_bounds__(BS) k_ir_inner_init(
    int64_t n, const double *r, const float *diag32, float *r32, float *x32,
    double *partials, const State *so, State *si)
{
    if (so->done)
        return;
    const double inv = 1.0 / so->res;
    double v[3] = {0.0, 0.0, 0.0}, o[3];
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS) {
        float ri = (float)(r[i] * inv);
        r32[i] = ri;
        x32[i] = 0.0f;
        float d = diag32[i];
        v[0] += (double)ri * (double)ri;
        if (d == 0.0f)
            v[2] += 1.0;
        else
            v[1] += (double)ri * (double)(ri / d);
    }
    if (reduce_last<3>(v, partials, &si->counter[0], o) && threadIdx.x == 0) {
        double res = sqrt(o[0]);
        si->nits = 0;
        si->res = si->r0 = res;
        si->tol = res * si->rtol;
        si->rz = o[1];
        si->rz_prev = 0.0;
        si->done = res > si->tol ? (o[2] != 0.0 ? ZERO_DIAG : RUNNING)
                                 : CONVERGED;
    }
}

// x += |r| * d  (FP64 correction)
__global__ void __launch_bounds__(BS) k_ir_correct(int64_t n, double *x,
                                                   const float *d,
                                                   State *so)
{
    if (so->done)
        return;
    const double s = so->res;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        x[i] += s * (double)d[i];
    if (blockIdx.x == 0 && threadIdx.x == 0)
        so->outer += 1;   // visible to the next kernel on the stream
}

__global__ void k_to_f32(int64_t n, const double *a, float *b)
{
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BS)
        b[i] = (float)a[i];
}

//------------------------------------------------------------- host side
struct Workspace {
    int64_t n = 0;
    double *r = nullptr, *p = nullptr, *Ap = nullptr, *partials = nullptr;
    float *r32 = nullptr, *p32 = nullptr, *Ap32 = nullptr, *x32 = nullptr;
    State *st = nullptr;      // [2]: outer/FP64, inner/FP32
};

constexpr int MAX_GRID = 256 * 8;   // partial-sum slots per reduced value

int grid_for(int64_t n)
{
    int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    int64_t g = (n + BS - 1) / BS;
    int64_t cap = std::min<int64_t>((int64_t)cus * 8, MAX_GRID);
    return (int)std::max<int64_t>(1, std::min(g, cap));
}

int grid_rows(int64_t n, int L)
{
    // each work-group covers a contiguous range of several row groups
    int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    int64_t groups = (n + BS / L - 1) / (BS / L);
    int64_t cap = std::min<int64_t>((int64_t)cus * 8, MAX_GRID);
    return (int)std::max<int64_t>(1, std::min(groups, cap));
}

Workspace *get_ws(const phg_hip_mat *A, int need32)
{
    static Workspace ws;     // one solve at a time per process (one stream)
    if (ws.n != A->n) {
        phg::dfree(ws.r); phg::dfree(ws.p); phg::dfree(ws.Ap);
        phg::dfree(ws.partials); phg::dfree(ws.st);
        phg::dfree(ws.r32); phg::dfree(ws.p32); phg::dfree(ws.Ap32);
        phg::dfree(ws.x32);
        ws = Workspace();
        size_t n8 = (size_t)A->n * sizeof(double);
        if (phg::dmalloc((void **)&ws.r, n8) || phg::dmalloc((void **)&ws.p, n8) ||
            phg::dmalloc((void **)&ws.Ap, n8) ||
            phg::dmalloc((void **)&ws.partials, 4 * MAX_GRID * sizeof(double)) ||
            phg::dmalloc((void **)&ws.st, 2 * sizeof(State)))
            return nullptr;
        if (hipMemsetAsync(ws.st, 0, 2 * sizeof(State), phg::ctx()->stream) != hipSuccess)
            return nullptr;
        ws.n = A->n;
    }
    if (need32 && ws.r32 == nullptr) {
        size_t n4 = (size_t)A->n * sizeof(float);
        if (phg::dmalloc((void **)&ws.r32, n4) || phg::dmalloc((void **)&ws.p32, n4) ||
            phg::dmalloc((void **)&ws.Ap32, n4) || phg::dmalloc((void **)&ws.x32, n4))
            return nullptr;
    }
    return &ws;
}

#define LANES_SWITCH(L, CALL)                                                 \
    switch (L) {                                                              \
    case 4: { constexpr int LL = 4; CALL; } break;                            \
    case 16: { constexpr int LL = 16; CALL; } break;                          \
    case 32: { constexpr int LL = 32; CALL; } break;                          \
    default: { constexpr int LL = 8; CALL; } break;                           \
    }

// one PCG iteration (3 kernels) on the given precision
template <typename T>
void launch_iteration(const phg_hip_mat *A, int L, const T *va, const T *diag,
                      T *x, T *r, T *p, T *Ap, double *partials, State *st,
                      hipStream_t s)
{
    int64_t n = A->n;
    int g = grid_for(n), gr = grid_rows(n, L);
    hipLaunchKernelGGL(k_pcg_p<T>, dim3(g), dim3(BS), 0, s, n, (const T *)r,
                       diag, p, (const State *)st);
    LANES_SWITCH(L, hipLaunchKernelGGL((k_pcg_spmv<LL, T>), dim3(gr), dim3(BS),
                                       0, s, n, (const int32_t *)A->d_rowptr,
                                       (const int32_t *)A->d_cols, va,
                                       (const T *)p, Ap, partials, st));
    hipLaunchKernelGGL(k_pcg_update<T>, dim3(g), dim3(BS), 0, s, n, x,
                       (const T *)p, r, (const T *)Ap, diag, partials, st);
}

int read_state(State *d, State *h, int count, hipStream_t s)
{
    PHG_HIP_CHECK(hipMemcpyAsync(h, d, count * sizeof(State),
                                 hipMemcpyDeviceToHost, s));
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    return PHG_HIP_OK;
}

int upload_state(State *d, const State &h, hipStream_t s)
{
    PHG_HIP_CHECK(hipMemcpyAsync(d, &h, sizeof(State), hipMemcpyHostToDevice, s));
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

}  // namespace

namespace phg {

int choose_spmv_shape(phg_hip_mat *A)
{
    // lanes per row from the mean row length: one or two passes per lane
    double mean = A->n > 0 ? (double)A->nnz / (double)A->n : 0.0;
    int L = mean <= 6 ? 4 : (mean <= 20 ? 8 : (mean <= 48 ? 16 : 32));
    A->spmv_lanes = L;
    A->spmv_lanes32 = L;
    return PHG_HIP_OK;
}

int ensure_fp32_copy(phg_hip_mat *A)
{
    if (A->d_vals32 != nullptr)
        return PHG_HIP_OK;
    hipStream_t s = ctx()->stream;
    int rc;
    if ((rc = dmalloc((void **)&A->d_vals32, (size_t)(A->nnz + 1) * sizeof(float))) ||
        (rc = dmalloc((void **)&A->d_diag32, (size_t)A->n * sizeof(float))))
        return rc;
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->nnz)), dim3(BS), 0, s, A->nnz,
                       (const double *)A->d_vals, A->d_vals32);
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->n)), dim3(BS), 0, s, A->n,
                       (const double *)A->d_diag, A->d_diag32);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    return PHG_HIP_OK;
}

int spmv64(const phg_hip_mat *A, const double *x, double *y, double alpha,
           double beta, hipStream_t s)
{
    int L = A->spmv_lanes;
    int gr = grid_rows(A->n, L);
    LANES_SWITCH(L, hipLaunchKernelGGL((k_spmv64<LL>), dim3(gr), dim3(BS), 0,
                                       s, A->n, (const int32_t *)A->d_rowptr,
                                       (const int32_t *)A->d_cols,
                                       (const double *)A->d_vals, x, y, alpha,
                                       beta));
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

}  // namespace phg

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
    unsigned *cnt = nullptr;
    double *out = nullptr;
    if ((rc = phg::dmalloc((void **)&part, (MAX_GRID + 2) * sizeof(double))))
        return rc;
    cnt = (unsigned *)(part + MAX_GRID);
    out = part + MAX_GRID + 1;
    PHG_HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned), s));
    int g = grid_for(x->n);
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(BS), 0, s, x->n,
                       (const double *)x->d, (const double *)y->d, part, cnt,
                       out);
    PHG_HIP_CHECK(hipGetLastError());
    rc = phg::download(res, out, sizeof(double));
    phg::dfree(part);
    return rc;
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

extern "C" int phg_hip_pcg(const phg_hip_mat *A, const phg_hip_vec *b,
                           phg_hip_vec *x, int32_t maxit, double rtol,
                           double btol, double atol, double *residual)
{
    int rc;
    if ((rc = check_solve_args(A, b, x)))
        return rc;
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
    if ((rc = upload_state(ws->st, h, s)))
        return rc;
    const int64_t n = A->n;
    const int L = A->spmv_lanes;
    int g = grid_for(n), gr = grid_rows(n, L);
    hipLaunchKernelGGL(k_pcg_init_norms<8>, dim3(g), dim3(BS), 0, s, n,
                       (const double *)b->d, (const double *)x->d, ws->partials,
                       ws->st);
    LANES_SWITCH(L, hipLaunchKernelGGL((k_pcg_init_res<LL>), dim3(gr),
                                       dim3(BS), 0, s, n,
                                       (const int32_t *)A->d_rowptr,
                                       (const int32_t *)A->d_cols,
                                       (const double *)A->d_vals,
                                       (const double *)A->d_diag,
                                       (const double *)b->d, x->d, ws->r,
                                       ws->partials, ws->st));
    PHG_HIP_CHECK(hipGetLastError());
    State *hm = (State *)phg::ctx()->host_mailbox;
    int chunk = 8;
    int64_t launched = 0;
    for (;;) {
        if ((rc = read_state(ws->st, hm, 1, s)))
            return rc;
        if (hm->done != RUNNING)
            break;
        // the device stops at maxit; a state that never finishes means the
        // kernels did not run
        if (launched > (int64_t)maxit + 256)
            return phg::set_error(PHG_HIP_E_HIP, "PCG state did not advance");
        launched += chunk;
        for (int k = 0; k < chunk; k++)
            launch_iteration<double>(A, L, A->d_vals, A->d_diag, x->d, ws->r,
                                     ws->p, ws->Ap, ws->partials, ws->st, s);
        PHG_HIP_CHECK(hipGetLastError());
        chunk = std::min(chunk * 2, 64);
    }
    if (residual)
        *residual = hm->res;
    if (hm->done == ZERO_DIAG)
        return phg::set_error(PHG_HIP_E_ZERO_DIAG,
                              "zero diagonal in Jacobi preconditioner");
    return hm->nits;
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
    if ((rc = phg::ensure_fp32_copy(A)))
        return rc;
    hipStream_t s = phg::ctx()->stream;
    Workspace *ws = get_ws(A, 1);
    if (ws == nullptr)
        return phg::set_error(PHG_HIP_E_NOMEM, "workspace allocation failed");
    State h[2];
    memset(h, 0, sizeof(h));
    h[0].rtol = rtol;
    h[0].btol = btol;
    h[0].atol = atol;
    h[0].maxit = maxit;
    h[1].rtol = inner_rtol;
    h[1].maxit = 100000;   // each FP32 solve stops on inner_rtol
    PHG_HIP_CHECK(hipMemcpyAsync(ws->st, h, 2 * sizeof(State),
                                 hipMemcpyHostToDevice, s));
    const int64_t n = A->n;
    const int L = A->spmv_lanes, L32 = A->spmv_lanes32;
    int g = grid_for(n), gr = grid_rows(n, L);
    State *so = ws->st, *si = ws->st + 1;
    State *hm = (State *)phg::ctx()->host_mailbox;
    int total_inner = 0, outer = 0;
    for (;;) {
        LANES_SWITCH(L, hipLaunchKernelGGL((k_ir_residual<LL>), dim3(gr),
                                           dim3(BS), 0, s, n,
                                           (const int32_t *)A->d_rowptr,
                                           (const int32_t *)A->d_cols,
                                           (const double *)A->d_vals,
                                           (const double *)b->d,
                                           (const double *)x->d, ws->r,
                                           ws->partials, so));
        hipLaunchKernelGGL(k_ir_inner_init, dim3(g), dim3(BS), 0, s, n,
                           (const double *)ws->r, (const float *)A->d_diag32,
                           ws->r32, ws->x32, ws->partials, (const State *)so,
                           si);
        PHG_HIP_CHECK(hipGetLastError());
        if ((rc = read_state(ws->st, hm, 2, s)))
            return rc;
        if (hm[0].done != RUNNING || outer >= maxit)
            break;
        if (hm[1].done == ZERO_DIAG)
            return phg::set_error(PHG_HIP_E_ZERO_DIAG,
                                  "zero diagonal in Jacobi preconditioner");
        int chunk = 16;
        int64_t launched = 0;
        for (;;) {
            if (launched > (int64_t)h[1].maxit + 256)
                return phg::set_error(PHG_HIP_E_HIP,
                                      "inner PCG state did not advance");
            launched += chunk;
            for (int k = 0; k < chunk; k++)
                launch_iteration<float>(A, L32, A->d_vals32, A->d_diag32,
                                        ws->x32, ws->r32, ws->p32, ws->Ap32,
                                        ws->partials, si, s);
            PHG_HIP_CHECK(hipGetLastError());
            if ((rc = read_state(si, hm + 1, 1, s)))
                return rc;
            if (hm[1].done != RUNNING)
                break;
            chunk = std::min(chunk * 2, 64);
        }
        total_inner += hm[1].nits;
        hipLaunchKernelGGL(k_ir_correct, dim3(g), dim3(BS), 0, s, n, x->d,
                           (const float *)ws->x32, so);
        PHG_HIP_CHECK(hipGetLastError());
        outer++;
    }
    if (inner_its)
        *inner_its = total_inner;
    if (residual)
        *residual = hm[0].res;
    return outer;
}
