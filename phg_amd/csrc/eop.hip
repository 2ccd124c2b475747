// Matrix-free FP32 operator of a scalar stiffness (+ mass) matrix, element
// products on the matrix cores, row sums in LDS.
//
// For an affine tetrahedron the stiffness of quad.c:5260
// (phgQuadGradBasAGradBas with a P0 coefficient) factors as
//     K_e = sum_{m=0..5} c_m(e) T^m,
//     c_m(e) = kappa_e vol_e (grad lambda_a . grad lambda_b),  m = (a <= b),
// over the independent barycentric coordinates a, b = 1..3, with reference
// tensors T^m fixed per order (assemble.hip computes both).  A mass term
// a phgQuadBasDotBas (quad.c:5582) is the seventh product, c_6 = a vol_e and
// T^6 the reference mass matrix, in slot 6 of the tiles.  Applying the
// operator to x is then, per element,
//     y_e[i] = sum_m c_m(e) (T^m x_e)[i],
// and T^m x_e for 16 elements at a time is a matrix product: rows (i, m) of
// the stacked T against the 16 gathered element vectors.  It runs as
// v_mfma_f32_16x16x32_bf16 with FP32 precision kept by splitting both
// operands into three bf16 parts (t = t0 + t1 + t2, x = x0 + x1 + x2, exact
// for FP32 x) and accumulating the six products whose order is at most
// 2^-16 relative (t0x0, t0x1, t1x0, t0x2, t1x1, t2x0) in FP32; the dropped
// terms are below 2^-24.  One 16-row output tile holds the slots m (8 rows,
// 6 used, 7 with a mass term) of two local indices i = 2t, 2t + 1 for 16 elements; the A
// fragments (T parts) of all tiles stay in LDS for the kernel, the B
// fragments (x parts) are written to LDS per 16 elements.
//
// P2 takes the products through the points of the rule the tensors are
// summed on (order 2, four points q, weights w_q):
//     U_e[q][b] = sum_j D_jb(q) x_e[j],   V_e[q][a] = w_q sum_b c_ab(e) U_e[q][b],
//     y_e[i]    = sum_{q,a} D_ia(q) V_e[q][a],
// D_ia(q) the gradient of phi_i in the independent barycentric coordinates
// at point q and c_ab the symmetric metric of the c_m -- the same sum in
// another order.  Both products are v_mfma_f32_16x16x4f32 of one 16 x 12
// tile (rows (q, a) against the local indices j, then rows i against (q, a)),
// six MFMA per 16 elements against fifteen for the stacked tensors, and the
// 3 x 3 metric between them is applied per lane.  With a mass term the third
// product c_6 T^6 x_e goes into the same accumulator.
//
// Element blocks.  The elements, in Hilbert-curve order of their centroids, are cut
// into blocks of at most EBLK elements touching at most KMAX rows.  A
// work-group takes a block at a time:
//   1. x of the block's touched rows into LDS (one gather per row);
//   2. the element products, each element-local result stored into an LDS
//      slot -- the slots of a row are consecutive and in element order;
//   3. one work-item per touched row sums its slots in that order.  A row
//      touched by this block only (interior) is written to y directly; a
//      row shared with other blocks writes a partial, summed in block order
//      by k_eop_fix or, in the solvers, by their update kernel.  Each block
//      also sets a slice of the Dirichlet rows (diag * x), so (x, y) is
//      complete when the block kernel ends.
// Every sum has a fixed order, so the operator is deterministic.  Per
// element the kernel streams 4 B per local index (touched-row index and slot,
// packed) and 24 B of coefficients, about 1/3 of the FP32 CSR entries the
// element contributes (P2: 64 B against ~240 B).
#include "phg_hip_internal.h"

#include <hipcub/hipcub.hpp>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <memory>
#include <numeric>
#include <thread>
#include <utility>
#include <vector>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int EB = 512;           // work-items per work-group (8 wavefronts)
constexpr int EW = EB / 64;
constexpr int NREF = 7;           // reference tensors: six stiffness, one mass
// coefficients per element: the six stiffness ones, then the mass term's
template <bool MASS> constexpr int kNc = MASS ? 7 : 6;
constexpr int FB = 256;           // work-items of the fix-up kernel
constexpr uint32_t KD_NONE = 0xffffffffu;

// Host set-up work split over up to 16 threads: f(lo, hi, t) for contiguous
// chunks [lo, hi) of [0, n), at least `grain` items each, t the chunk
// number.  A thread that cannot be started runs its chunk on the caller.  An
// allocation that fails in any chunk is reported to the caller as
// std::bad_alloc once every chunk has ended.
inline int host_threads(int64_t n, int64_t grain = 1024)
{
    const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
    return (int)std::max<int64_t>(1, std::min<int64_t>({16, hw, n / grain}));
}

template <typename F>
void host_parallel(int64_t n, F &&f, int64_t grain = 1024)
{
    const int T = host_threads(n, grain);
    std::vector<std::thread> th;
    std::atomic<bool> failed{false};
    auto run = [&f, &failed](int64_t lo, int64_t hi, int t) {
        try {
            f(lo, hi, t);
        }
        catch (...) {
            failed = true;
        }
    };
    for (int t = 1; t < T; t++) {
        const int64_t lo = n * t / T, hi = n * (t + 1) / T;
        try {
            th.emplace_back([&run, lo, hi, t] { run(lo, hi, t); });
        }
        catch (...) {
            run(lo, hi, t);
        }
    }
    run(0, n / T, 0);
    for (auto &x : th)
        x.join();
    if (failed)
        throw std::bad_alloc();
}

// Allocator whose value-initialisation is a no-op: the large set-up arrays
// are first touched (so paged in) by the host threads that fill them.
template <typename T>
struct NoInit : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <typename U>
    NoInit(const NoInit<U> &) {}
    template <typename U>
    void construct(U *) noexcept {}
    template <typename U, typename... Args>
    void construct(U *q, Args &&...args)
    {
        ::new ((void *)q) U(std::forward<Args>(args)...);
    }
};
template <typename T>
using HostVec = std::vector<T, NoInit<T>>;

template <int P> struct EopShape {
    static constexpr int NB = P == 1 ? 4 : (P == 2 ? 10 : 20);
    static constexpr int NT = NB / 2;                 // output tiles
    static constexpr int KS = (6 * NB + 31) / 32;     // K steps of 32
    static constexpr int KP = KS * 32 + 8;            // B row stride (bank skew)
    // per lane: x of NJ local indices j, results of NO local indices i;
    // their 16-bit indices packed into a chunk of W dwords
    static constexpr int NJ = (NB + 3) / 4;
    static constexpr int NO = (NT + 1) / 2;
    static constexpr int W = (NJ + NO + 1) / 2;
    // elements and touched rows per block; slots per block (< 2^16)
    static constexpr int EBLK = P == 1 ? 2048 : (P == 2 ? 1024 : 512);
    static constexpr int KMAX = P == 1 ? 1024 : (P == 2 ? 2048 : 3072);
    static constexpr int SMAX = EBLK * NB;
    static_assert(SMAX < 65535 && KMAX < 65535, "16-bit slot and row indices");
};

// A fragment of (tile t, K step s) for lane l: row R = l % 16 is slot
// m = R % 8 of local index i = 2 t + R / 8; K indices 32 s + 8 (l / 16) + u,
// u = 0..7; K index k = q NB + j is product q of the split, column j of
// T^m (row i).
constexpr int kTPart[6] = {0, 0, 1, 0, 1, 2};
constexpr int kXPart[6] = {0, 1, 0, 2, 1, 0};

// element tiles of 16 a wavefront takes per round: their gathers are
// issued together and every A fragment read from LDS feeds ET products
template <int P> constexpr int kEt = P == 1 ? 2 : 1;

// P3 runs the bf16 split products above; P1 / P2 run FP32 MFMA
// (v_mfma_f32_16x16x4f32): K = NB local indices in steps of 4, the A
// fragments (one T value per lane and step) in registers and the B
// fragments straight from the lane's gathered x, no split and no LDS
// staging.  Lane layout: A[l % 16][l / 16], B[l / 16][l % 16], D rows
// 4 (l / 16) + r of column l % 16.
template <int P> constexpr bool kSplit = P == 3;

// P2: products through the rule's points (see above).  Lane l = 16 g + n of
// the first product holds U rows 4 g + r = (q = g, a = r), r < 3, of element
// n -- the point and all three directions the metric needs -- and hands
// V[q = g][a = s] to K step s of the second product as its B fragment.  The
// output rows 4 g + r are the local indices kOutRow2(4 g + r): three per lane
// group for g = 0, 1, two for g = 2, 3.
template <int P> constexpr bool kPts = P == 2;
constexpr int kNp2 = 4;               // points of the order-2 rule

__host__ __device__ constexpr int kOutRow2(int row)
{
    return row / 4 < 2 ? (row % 4 < 3 ? 3 * (row / 4) + row % 4 : -1)
                       : (row % 4 < 2 ? 6 + 2 * (row / 4 - 2) + row % 4 : -1);
}

struct BlockView {
    int64_t nblk;
    const int32_t *eb_elem;   // [nblk + 1] element ranges
    const int32_t *eb_row;    // [nblk + 1] touched-row ranges
    const int32_t *eb_nint;   // [nblk] interior rows (first in the list)
    const int32_t *rows;      // [nent] solver rows
    const uint16_t *rs;       // [nent + nblk] slot starts, block b at eb_row[b] + b
    const int32_t *ppos;      // [nent] partial position of a shared row
    const uint32_t *idx;      // [nelem][4][W] index chunks of the four lane groups
    const float *coef;        // [nelem][6 or 7]
    const void *tiles;        // A fragments: bf16 (split path) or FP32
    float *part;              // partials of shared rows
    const int32_t *drows;     // [ndir] Dirichlet rows, spread evenly over the blocks
    int64_t ndir;
    const float *diag;        // FP32 diagonal (Dirichlet rows: y = diag x)
    float msum[16];           // mass term: row sums of T^6 (P1 / P3 per tile t, P2 per row)
};

// work-group total of one double per work-item, to out[blockIdx.x]
template <int NTH>
__device__ inline void group_total(double v, double *out)
{
    __shared__ double sh[NTH / 64];
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NTH / 64; k++)
            s += sh[k];
        out[blockIdx.x] = s;
    }
}

// the wavefront's own LDS rows written by some lanes and read by others:
// LDS operations of one wavefront take effect in issue order, so only the
// compiler has to be kept from moving them across this point
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// element of index idx of an array under 4 GiB: a 32-bit byte offset from
// the base, so the load takes the base in scalar registers
template <typename T>
__device__ inline T ld32(const T *base, uint32_t idx)
{
    return *(const T *)((const char *)base + idx * (uint32_t)sizeof(T));
}

// Lane l = 16 g + n of a wavefront works on element n of a tile.  It
// gathers x of the local indices j = NJ g + u (split path; the B fragment K
// positions of j are filled through LDS) or j = 4 u + g (FP32 path: K step
// u), u < NJ, and stores the results of the local
// indices i = 2 t + g / 2 with t = g mod 2 (the D rows it holds after the
// contraction), t = 2 o + g % 2 (o < NO) -- on the P2 point-product path
// the local indices kOutRow2(4 g + o) instead.  Its chunk of the element's index
// record holds, as 16-bit values, the block-local row of each j and the
// block-local slot of each i (0xffff: Dirichlet / none).
//
// One round = ET element tiles.  The loads are branch-free (element clamped
// into the block) and hold raw values; elements past the block are masked
// where the results are stored, so the loads of a round stay in flight
// while the previous round computes.
template <int ET, int W, int NCF> struct Round {
    uint32_t ch[ET][W];
    float cf[ET][NCF];
};

// coefficients a lane holds: four (the tiles of its row group), or all of
// its element's on the point-product path
template <int P, bool MASS> constexpr int kNcf = kPts<P> ? kNc<MASS> : 4;

template <int P, int W, int ET, bool MASS>
__device__ inline void load_round(const BlockView &m, int32_t e0, int32_t ne, int w, int n,
                                  int g, int rd, Round<ET, W, kNcf<P, MASS>> &r)
{
    // 32-bit offsets: nelem * nb < 2^30 is checked at set-up
#pragma unroll
    for (int et = 0; et < ET; et++) {
        const int el = ((rd * EW + w) * ET + et) * 16 + n;
        const uint32_t e = (uint32_t)(e0 + min(el, ne - 1));
        // P2 lane groups 2, 3 hold two dwords (the third load repeats the
        // second; compute() spreads them into the three-dword form)
        constexpr bool P2 = kPts<P>;
        const uint32_t base = e * (uint32_t)phg::eop_record_dwords(P) +
                              (uint32_t)(P2 ? (g < 2 ? 3 * g : 2 * g + 2) : g * W);
#pragma unroll
        for (int u = 0; u < W; u++)
            r.ch[et][u] = ld32(m.idx, base + (uint32_t)(P2 && g >= 2 ? min(u, 1) : u));
        // coefficients m = 4 (g % 2) + q; for odd g the slots past the
        // element's last coefficient (7, and 6 without the mass term) read
        // the next element's (finite: the array has two entries of padding)
        // and meet all-zero rows of the T tiles.  Point-product path: all.
        if constexpr (kPts<P>) {
#pragma unroll
            for (int q = 0; q < kNc<MASS>; q++)
                r.cf[et][q] = ld32(m.coef, e * kNc<MASS> + q);
        }
        else {
#pragma unroll
            for (int q = 0; q < 4; q++)
                r.cf[et][q] = ld32(m.coef, e * kNc<MASS> + 4 * (g & 1) + q);
        }
    }
}

// 16-bit entry h of a lane's index chunk
template <int W>
__device__ inline uint32_t chunk_half(const uint32_t (&ch)[W], int h)
{
    return (ch[h >> 1] >> (16 * (h & 1))) & 0xffffu;
}

// x of a lane's local indices, minus x_0 of its element (local index 0,
// held by lane n of the wavefront), and x_0.  Every stiffness tensor T^m
// annihilates constants, so T^m x_e = T^m (x_e - x_0 1) exactly; in FP32 the
// products then carry rounding errors relative to the variation of x over
// the element rather than to x itself -- the local differences of a smooth x
// are exact (Sterbenz) and the Krylov iterations converge as on the
// assembled FP32 matrix.  The mass tensor does not: its product takes
// x_0 (T^6 1)_i back after the contraction.  Dirichlet and absent local
// indices (x = 0, row KMAX of xs) and local indices past NB (zero rows of T)
// take -x_0 alike.
template <int KMAX, int NJ, int ET, int W>
__device__ inline void load_relative(const float *xs, const uint32_t (&ch)[ET][W], int n,
                                     float (&xv)[ET][NJ], float (&x0)[ET])
{
#pragma unroll
    for (int et = 0; et < ET; et++) {
#pragma unroll
        for (int u = 0; u < NJ; u++)
            xv[et][u] = xs[min(chunk_half(ch[et], u), (uint32_t)KMAX)];
        x0[et] = __shfl(xv[et][0], n, 64);
#pragma unroll
        for (int u = 0; u < NJ; u++)
            xv[et][u] -= x0[et];
    }
}

// wavefronts per SIMD the block kernel is compiled for: three work-groups
// per CU for P2, two for P1 (LDS and VGPRs both fit), one for P3
template <int P> constexpr int kWaves = P == 3 ? 2 : (P == 2 ? 6 : 4);

// y = A x over the element blocks: interior and Dirichlet rows of y,
// partials of shared rows, the work-group's part of (x, y) over all rows
// in part_out
template <int P, bool MASS>
__global__ void __launch_bounds__(EB, kWaves<P>)
k_eop_block(BlockView m, const float *__restrict__ x, float *__restrict__ y,
            double *part_out, const int32_t *done, const int32_t *mode, int32_t want)
{
    using S = EopShape<P>;
    constexpr int NB = S::NB, NT = S::NT, KS = S::KS, KP = S::KP, ET = kEt<P>;
    constexpr int NJ = S::NJ, NO = S::NO, W = S::W;
    constexpr int RPT = (S::KMAX + EB - 1) / EB;   // touched rows per work-item
    if (done != nullptr && (*done != 0 || *mode != want))
        return;
    constexpr bool SPLIT = kSplit<P>, PTS = kPts<P>;
    constexpr int NCF = kNcf<P, MASS>;
    // point-product A values: first product (NJ steps), second (3 steps),
    // mass tensor (NJ steps)
    constexpr int NAP = 2 * NJ + 3;
    __shared__ uint4 at[SPLIT ? NT * KS * 64 : 1];
    __shared__ __align__(16) __bf16 bt[SPLIT ? EW : 1][16][KP];
    __shared__ float xs[S::KMAX + 1];
    __shared__ float slot[S::SMAX + 1];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int n = l & 15, g = l >> 4;
    float af[SPLIT || PTS ? 1 : NT][SPLIT || PTS ? 1 : NJ];
    float ap[PTS ? NAP : 1];
    if constexpr (SPLIT) {
        for (int t = tid; t < NT * KS * 64; t += EB)
            at[t] = ((const uint4 *)m.tiles)[t];
        // K padding of the B rows stays zero
        constexpr int PADW = KP - 6 * NB;
        for (int t = l; t < 16 * PADW; t += 64)
            bt[w][t / PADW][6 * NB + t % PADW] = (__bf16)0.0f;
    }
    else if constexpr (PTS) {
#pragma unroll
        for (int k = 0; k < (MASS ? NAP : NJ + 3); k++)
            ap[k] = ((const float *)m.tiles)[k * 64 + l];
    }
    else {
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int u = 0; u < NJ; u++)
                af[t][u] = ((const float *)m.tiles)[(t * NJ + u) * 64 + l];
    }
    if (tid == 0)
        xs[S::KMAX] = 0.0f;     // x of Dirichlet and absent local indices
    double dot = 0.0;
    int64_t first, end, step;
    phg::xcd_sweep(m.nblk, first, end, step);
    // block header; touched-row loads of a block (index clamped, the row
    // tables carry one entry of padding; raw values, masked where used)
    auto header = [&](int64_t bb, int32_t &k0, int32_t &nk, int32_t &nint, int32_t &e0,
                      int32_t &ne) {
        k0 = m.eb_row[bb];
        nk = m.eb_row[bb + 1] - k0;
        nint = m.eb_nint[bb];
        e0 = m.eb_elem[bb];
        ne = m.eb_elem[bb + 1] - e0;
    };
    auto row_loads = [&](int64_t bb, int32_t k0, int32_t nk, int32_t (&rw)[RPT],
                         uint32_t (&qr)[RPT]) {
        const int kc = max(nk - 1, 0);
        const uint16_t *rs = m.rs + k0 + bb;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const int k = min(tid + EB * i, kc);
            rw[i] = m.rows[k0 + k];
            qr[i] = (uint32_t)rs[k] | (uint32_t)rs[k + 1] << 16;
        }
    };
    if (first < end) {
        // software pipeline over the work-group's blocks: while block b
        // computes, the touched rows, x values and first element round of
        // the next block are already loading
        int64_t b = first;
        int32_t k0, nk, nint, e0, ne;
        header(b, k0, nk, nint, e0, ne);
        int32_t rw[RPT];
        uint32_t qr[RPT];
        float xr[RPT];
        row_loads(b, k0, nk, rw, qr);
#pragma unroll
        for (int i = 0; i < RPT; i++)
            xr[i] = x[rw[i]];
        Round<ET, W, NCF> ra, rb;
        load_round<P, W, ET, MASS>(m, e0, ne, w, n, g, 0, ra);
        // one round: x from the block's rows, B fragments, products, slots
        auto compute = [&](int rd, Round<ET, W, NCF> &cur) {
            float yv[ET][NO];
            // x of the lane's local indices (split path j = NJ g + u, FP32
            // path j = 4 u + g; Dirichlet and none: row KMAX of xs, which
            // holds 0), relative to x_0
            float xv[ET][NJ], xz[ET];
            if constexpr (PTS) {
                // lane groups 2, 3: [j0 j1][i0 i1] -> [j0 j1][none i0][i1 none]
#pragma unroll
                for (int et = 0; et < ET; et++) {
                    const uint32_t c1 = cur.ch[et][1];
                    cur.ch[et][1] = g >= 2 ? 0xffffu | c1 << 16 : c1;
                    cur.ch[et][2] = g >= 2 ? c1 >> 16 | 0xffff0000u : cur.ch[et][2];
                }
            }
            load_relative<S::KMAX, NJ, ET, W>(xs, cur.ch, n, xv, xz);
            // mass term: the lanes of slots 4..7 (odd g, c_6 their third
            // coefficient) add c_6 x_0 (T^6 1)_i to their part of row i
            float mc[ET];
#pragma unroll
            for (int et = 0; et < ET; et++)
                mc[et] = PTS ? 0.0f : ((g & 1) ? cur.cf[et][2] * xz[et] : 0.0f);
            // contract tile t's products with c_m, then add the two row
            // groups of local index 2 t + g / 2 (lanes g and g ^ 1)
            auto contract = [&](int et, int t, const f32x4 &a) {
                float v = cur.cf[et][0] * a[0] + cur.cf[et][1] * a[1] + cur.cf[et][2] * a[2] +
                          cur.cf[et][3] * a[3];
                if constexpr (MASS)
                    v += mc[et] * m.msum[t];
                v += __shfl_xor(v, 16, 64);
                if ((t & 1) == (g & 1))
                    yv[et][t >> 1] = v;
            };
            if constexpr (PTS) {
#pragma unroll
                for (int et = 0; et < ET; et++) {
                    const float *c = cur.cf[et];
                    // U rows (q = g, a = r) of element n
                    f32x4 U = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int u = 0; u < NJ; u++)
                        U = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[u], xv[et][u], U, 0, 0, 0);
                    // the metric at the lane's point; w_q is in the second
                    // product's A values
                    const float v[3] = {c[0] * U[0] + c[1] * U[1] + c[2] * U[2],
                                        c[1] * U[0] + c[3] * U[1] + c[4] * U[2],
                                        c[2] * U[0] + c[4] * U[1] + c[5] * U[2]};
                    f32x4 Y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int s2 = 0; s2 < 3; s2++)
                        Y = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[NJ + s2], v[s2], Y, 0, 0, 0);
                    if constexpr (MASS) {
                        // c_6 T^6 (x_e - x_0) in the same accumulator, then
                        // c_6 x_0 (T^6 1)_i
#pragma unroll
                        for (int u = 0; u < NJ; u++)
                            Y = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[NJ + 3 + u],
                                                                     c[6] * xv[et][u], Y, 0, 0, 0);
                        const float c0 = c[6] * xz[et];
#pragma unroll
                        for (int o = 0; o < NO; o++)
                            yv[et][o] = Y[o] + c0 * m.msum[4 * g + o];
                    }
                    else {
#pragma unroll
                        for (int o = 0; o < NO; o++)
                            yv[et][o] = Y[o];
                    }
                }
            }
            else if constexpr (SPLIT) {
                // split into bf16 parts through the wavefront's LDS rows, one
                // tile at a time, into B fragments
                bf16x8 bf[ET][KS];
#pragma unroll
                for (int et = 0; et < ET; et++) {
                    wave_lds_sync();    // the previous fragments are read
#pragma unroll
                    for (int u = 0; u < NJ; u++) {
                        const int j = NJ * g + u;
                        const float v = xv[et][u];
                        const __bf16 x0 = (__bf16)v;
                        const float r1 = v - (float)x0;
                        const __bf16 x1 = (__bf16)r1;
                        const __bf16 x2 = (__bf16)(r1 - (float)x1);
                        const __bf16 xp[3] = {x0, x1, x2};
                        // j past NB (its x is 0) goes to the zero K padding
#pragma unroll
                        for (int q = 0; q < 6; q++)
                            bt[w][n][j < NB ? q * NB + j : 6 * NB] = xp[kXPart[q]];
                    }
                    wave_lds_sync();
#pragma unroll
                    for (int s2 = 0; s2 < KS; s2++)
                        bf[et][s2] = *(const bf16x8 *)&bt[w][n][32 * s2 + 8 * g];
                }
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    f32x4 acc[ET];
#pragma unroll
                    for (int et = 0; et < ET; et++)
                        acc[et] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int s2 = 0; s2 < KS; s2++) {
                        const bf16x8 a = __builtin_bit_cast(bf16x8, at[(t * KS + s2) * 64 + l]);
#pragma unroll
                        for (int et = 0; et < ET; et++)
                            acc[et] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bf[et][s2],
                                                                              acc[et], 0, 0, 0);
                    }
#pragma unroll
                    for (int et = 0; et < ET; et++)
                        contract(et, t, acc[et]);
                }
            }
            else {
                // the relative x are the B fragments
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    f32x4 acc[ET];
#pragma unroll
                    for (int et = 0; et < ET; et++)
                        acc[et] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int u = 0; u < NJ; u++)
#pragma unroll
                        for (int et = 0; et < ET; et++)
                            acc[et] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[t][u], xv[et][u],
                                                                           acc[et], 0, 0, 0);
#pragma unroll
                    for (int et = 0; et < ET; et++)
                        contract(et, t, acc[et]);
                }
            }
            // element results into their slots; slot SMAX takes the
            // results nobody reads (none, Dirichlet, elements past the block)
#pragma unroll
            for (int et = 0; et < ET; et++) {
                const bool ok = ((rd * EW + w) * ET + et) * 16 + n < ne;
#pragma unroll
                for (int o = 0; o < NO; o++) {
                    const uint32_t sl = ok ? chunk_half(cur.ch[et], NJ + o) : 0xffffu;
                    slot[min(sl, (uint32_t)S::SMAX)] = yv[et][o];
                }
            }
        };
        while (true) {
            const int64_t bn = b + step;
            const bool more = bn < end;
            // the next block, or this one again past the last (its loads
            // are then harmless and unused)
            const int64_t bq = more ? bn : b;
            int32_t k0n, nkn, nintn, e0n, nen;
            header(bq, k0n, nkn, nintn, e0n, nen);
            uint32_t qq[RPT];
            __syncthreads();    // the previous block's xs and slot reads are done
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const bool in = tid + EB * i < nk;
                if (in)
                    xs[tid + EB * i] = xr[i];
                qq[i] = in ? qr[i] : 0u;
            }
            __syncthreads();
            int32_t rwn[RPT];
            uint32_t qrn[RPT];
            row_loads(bq, k0n, nkn, rwn, qrn);
            // loads this block needs after its rounds, issued now so that no
            // dependent load waits at the block's end: the partial positions
            // of its rows and its first EB Dirichlet rows (their x and
            // diagonal follow after the first round)
            const int kc = max(nk - 1, 0);
            int32_t opp[RPT];
#pragma unroll
            for (int i = 0; i < RPT; i++)
                opp[i] = m.ppos[k0 + min(tid + EB * i, kc)];
            const int64_t d0 = m.ndir * b / m.nblk, d1 = m.ndir * (b + 1) / m.nblk;
            const bool dhas = d0 + tid < d1;
            const int32_t dr = dhas ? m.drows[d0 + tid] : 0;
            float dxv = 0.0f, ddv = 0.0f;
            // rounds in pairs, the loads of round rd + 1 issued before round
            // rd computes; the round data never moves between registers (a
            // move would wait for the loads in flight).  Loads are
            // unconditional (a round past the block reads its last
            // element, an odd round count gets an empty last round) and the
            // last load of a block is the next block's first round
            const int ntile = (ne + 15) / 16;
            const int nround = (ntile + EW * ET - 1) / (EW * ET);
            for (int rd = 0; rd < nround; rd += 2) {
                load_round<P, W, ET, MASS>(m, e0, ne, w, n, g, rd + 1, rb);
                compute(rd, ra);
                if (rd == 0 && dhas) {
                    dxv = x[dr];
                    ddv = m.diag[dr];
                }
                const bool last = rd + 2 >= nround;
                load_round<P, W, ET, MASS>(m, last ? e0n : e0, last ? nen : ne, w, n, g,
                                  last ? 0 : rd + 2, ra);
                compute(rd + 1, rb);
            }
            // the next block's x
            float xrn[RPT];
#pragma unroll
            for (int i = 0; i < RPT; i++)
                xrn[i] = x[rwn[i]];
            __syncthreads();
            // row sums, slots in element order
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const int k = tid + EB * i;
                // four interleaved partial sums, combined in a fixed order
                const int q0 = qq[i] & 0xffffu, q1 = qq[i] >> 16;
                float s4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                int q = q0;
                for (; q + 4 <= q1; q += 4)
#pragma unroll
                    for (int u = 0; u < 4; u++)
                        s4[u] += slot[q + u];
                for (int u = 0; q < q1; q++, u++)
                    s4[u] += slot[q];
                const float sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
                // (x, y) of a shared row is the sum over its blocks of
                // x times the block's partial
                if (k < nk)
                    dot += (double)xs[k] * (double)sum;
                if (k < nint)
                    y[rw[i]] = sum;
                else if (k < nk)
                    m.part[opp[i]] = sum;
            }
            // this block's share of the Dirichlet rows: y = diag x
            {
                if (dhas) {
                    const float v = ddv * dxv;
                    y[dr] = v;
                    dot += (double)dxv * (double)v;
                }
                for (int64_t t = d0 + EB + tid; t < d1; t += EB) {
                    const int32_t r = m.drows[t];
                    const float xv = x[r], v = m.diag[r] * xv;
                    y[r] = v;
                    dot += (double)xv * (double)v;
                }
            }
            if (!more)
                break;
            b = bn;
            k0 = k0n;
            nk = nkn;
            nint = nintn;
            e0 = e0n;
            ne = nen;
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                rw[i] = rwn[i];
                qr[i] = qrn[i];
                xr[i] = xrn[i];
            }
        }
    }
    group_total<EB>(dot, part_out);
}

// shared rows: the sum of their partials in block order (the solvers do
// this in their update kernel instead when the per-row codes exist).  In a
// solver slot (done set) it returns at once, as the block kernel does, once
// the solve has finished or the slot is not an iteration slot.
__global__ void __launch_bounds__(FB)
k_eop_fix(int64_t nfx, const int32_t *__restrict__ rows, const int32_t *__restrict__ off,
          const float *__restrict__ part, float *__restrict__ y, const int32_t *done,
          const int32_t *mode, int32_t want)
{
    if (done != nullptr && (*done != 0 || *mode != want))
        return;
    for (int64_t t = (int64_t)blockIdx.x * FB + threadIdx.x; t < nfx;
         t += (int64_t)gridDim.x * FB) {
        float sum = 0.0f;
        const int32_t q1 = off[t + 1];
        for (int32_t q = off[t]; q < q1; q++)
            sum += part[q];
        y[rows[t]] = sum;
    }
}

// host bf16 (round to nearest even) of a float
uint16_t bf16_bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
double bf16_value(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return (double)f;
}

// A fragments [nb / 2][ks][64][8] of the stacked split T (see above); slot
// m = 6 holds the mass tensor when mass is set, else zeros
void build_tiles(int p, int nb, int ks, bool mass, std::vector<uint16_t> &out)
{
    const int nm = mass ? 7 : 6;
    std::vector<double> T((size_t)NREF * nb * nb);
    phg::eop_reference_tensors(p, T.data());
    // three-part split of every T entry
    std::vector<uint16_t> part((size_t)3 * T.size());
    for (size_t k = 0; k < T.size(); k++) {
        double r = T[k];
        for (int q = 0; q < 3; q++) {
            const uint16_t h = bf16_bits((float)r);
            part[(size_t)q * T.size() + k] = h;
            r -= bf16_value(h);
        }
    }
    const int nt = nb / 2;
    out.assign((size_t)nt * ks * 64 * 8, 0);
    for (int t = 0; t < nt; t++)
        for (int s = 0; s < ks; s++)
            for (int l = 0; l < 64; l++)
                for (int u = 0; u < 8; u++) {
                    const int R = l % 16, m = R % 8, i = 2 * t + R / 8;
                    const int k = 32 * s + 8 * (l / 16) + u;
                    const int q = k / nb, j = k % nb;
                    uint16_t v = 0;
                    if (m < nm && q < 6)
                        v = part[(size_t)kTPart[q] * T.size() +
                                 ((size_t)m * nb + i) * nb + j];
                    out[(((size_t)t * ks + s) * 64 + l) * 8 + u] = v;
                }
}

// FP32 A values [nb / 2][nj][64] of the FP32 path: lane l of tile t and
// K step u holds T^m[i][j], row R = l % 16 -> m = R % 8, i = 2 t + R / 8,
// j = 4 u + l / 16 (0 for j >= nb, for m = 7, and for m = 6 without mass)
void build_tiles_f32(int p, int nb, bool mass, std::vector<float> &out)
{
    const int nm = mass ? 7 : 6;
    std::vector<double> T((size_t)NREF * nb * nb);
    phg::eop_reference_tensors(p, T.data());
    const int nt = nb / 2, nj = (nb + 3) / 4;
    out.assign((size_t)nt * nj * 64, 0.0f);
    for (int t = 0; t < nt; t++)
        for (int u = 0; u < nj; u++)
            for (int l = 0; l < 64; l++) {
                const int R = l % 16, m = R % 8, i = 2 * t + R / 8, j = 4 * u + l / 16;
                if (m < nm && j < nb)
                    out[((size_t)t * nj + u) * 64 + l] = (float)T[((size_t)m * nb + i) * nb + j];
            }
}

// FP32 A values [2 nj + 3][64] of the P2 point-product path (lane l):
//   K step u < nj of the first product:  row R = l % 16 -> (q, a) = (R / 4,
//     R % 4), column j = 4 u + l / 16: D_ja(q) (0 for a = 3, j >= nb);
//   K step s < 3 of the second: row R -> local index i = kOutRow2(R),
//     column (q = l / 16, a = s): w_q D_ia(q) (0 where i < 0);
//   K step u < nj of the mass product: T^6[kOutRow2(R)][4 u + l / 16].
// Returns false unless the rule has kNp2 points.
bool build_tiles_pts(int p, int nb, std::vector<float> &out, float *msum)
{
    double w[phg::NPMAX_EOP], D[(size_t)phg::NPMAX_EOP * 20 * 3];
    if (phg::eop_rule_gradients(p, w, D) != kNp2)
        return false;
    std::vector<double> T((size_t)NREF * nb * nb);
    phg::eop_reference_tensors(p, T.data());
    const double *M = T.data() + (size_t)6 * nb * nb;
    const int nj = (nb + 3) / 4;
    out.assign((size_t)(2 * nj + 3) * 64, 0.0f);
    for (int l = 0; l < 64; l++) {
        const int R = l % 16, i = kOutRow2(R);
        for (int u = 0; u < nj; u++) {
            const int q = R / 4, a = R % 4, j = 4 * u + l / 16;
            if (a < 3 && j < nb)
                out[(size_t)u * 64 + l] = (float)D[((size_t)q * nb + j) * 3 + a];
            if (i >= 0 && j < nb)
                out[(size_t)(nj + 3 + u) * 64 + l] = (float)M[(size_t)i * nb + j];
        }
        for (int s = 0; s < 3; s++) {
            const int q = l / 16;
            if (i >= 0)
                out[(size_t)(nj + s) * 64 + l] = (float)(w[q] * D[((size_t)q * nb + i) * 3 + s]);
        }
    }
    for (int R = 0; R < 16; R++) {
        const int i = kOutRow2(R);
        double r = 0.0;
        for (int j = 0; i >= 0 && j < nb; j++)
            r += M[(size_t)i * nb + j];
        msum[R] = (float)r;
    }
    return true;
}

template <int P, bool MASS>
int grid_block()
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                     (const void *)k_eop_block<P, MASS>,
                                                     EB, 0) != hipSuccess ||
        per_cu <= 0) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    return per_cu * cus;
}

template <int P>
void launch_block(const EopData *eo, const float *x, const float *diag, float *y,
                  double *part_out, const int32_t *done, const int32_t *mode, int32_t want,
                  hipStream_t s)
{
    BlockView m{eo->nblk, eo->eb_elem, eo->eb_row, eo->eb_nint, eo->rows, eo->rs,
                eo->ppos, eo->idx, eo->coef, (const void *)eo->tiles, eo->part,
                eo->fx_rows + eo->nfx, eo->ndir, diag, {}};
    if (eo->mass) {
        for (int t = 0; t < 16; t++)
            m.msum[t] = eo->msum[t];
        hipLaunchKernelGGL((k_eop_block<P, true>), dim3(eo->grid_b), dim3(EB), 0, s, m, x, y,
                           part_out, done, mode, want);
    }
    else
        hipLaunchKernelGGL((k_eop_block<P, false>), dim3(eo->grid_b), dim3(EB), 0, s, m, x,
                           y, part_out, done, mode, want);
}

}  // namespace

namespace phg {

void free_eop(phg_hip_mat *A)
{
    EopData *eo = A->eop;
    if (eo == nullptr)
        return;
    dfree(eo->coef);
    dfree(eo->part);
    delete eo;          // drops this matrix's reference to the layout
    A->eop = nullptr;
}

void eop_apply(const EopData *eo, const float *x, const float *diag, float *y,
               double *part_out, const int32_t *done, const int32_t *mode, int32_t want,
               hipStream_t s, int which, bool fixup)
{
    if (which != 2) {
        if (eo->p == 1)
            launch_block<1>(eo, x, diag, y, part_out, done, mode, want, s);
        else if (eo->p == 2)
            launch_block<2>(eo, x, diag, y, part_out, done, mode, want, s);
        else
            launch_block<3>(eo, x, diag, y, part_out, done, mode, want, s);
    }
    if (which != 1 && fixup && eo->nfx > 0)
        hipLaunchKernelGGL(k_eop_fix, dim3(eo->grid_f), dim3(FB), 0, s, eo->nfx,
                           (const int32_t *)eo->fx_rows, (const int32_t *)eo->fx_off,
                           (const float *)eo->part, y, done, mode, want);
}

}  // namespace phg

EopLayout::~EopLayout()
{
    phg::dfree(eb_elem);
    phg::dfree(eb_row);
    phg::dfree(eb_nint);
    phg::dfree(rows);
    phg::dfree(rs);
    phg::dfree(ppos);
    phg::dfree(idx);
    phg::dfree(tiles);
    phg::dfree(tiles_mass);
    phg::dfree(fx_rows);
    phg::dfree(fx_off);
    phg::dfree(fxc);
    phg::dfree(eperm);
}

// Hilbert key of each element's centroid (0.25 times the sum of its vertex
// coordinates in vertex order), quantised to 21 bits per axis: the rounding
// steps are spelled out so the key does not depend on contraction
__global__ void k_elem_morton(int64_t ne, const double *__restrict__ xyz,
                              const int32_t *__restrict__ ev, double lo0, double lo1,
                              double lo2, double sc, uint64_t *key, int32_t *idx)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    const double lo[3] = {lo0, lo1, lo2};
    uint64_t q[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double m = 0.0;
#pragma unroll
        for (int v = 0; v < 4; v++)
            m = __dadd_rn(m, xyz[3 * (int64_t)ev[4 * e + v] + c]);
        const double f = __dmul_rn(__dsub_rn(__dmul_rn(0.25, m), lo[c]), sc);
        q[c] = (uint64_t)fmin(fmax(f, 0.0), 2097151.0);
    }
    key[e] = phg::hilbert3((uint32_t)q[0], (uint32_t)q[1], (uint32_t)q[2]);
    idx[e] = (int32_t)e;
}

namespace {

//==================================================== layout on the device ====
// The element-operator layout built on the device for blocks of EBLK
// consecutive elements (in Hilbert-curve order), when every such block touches at
// most KMAX rows -- the blocks the host's greedy split makes on such meshes,
// without its cuts.  Block tables, slot order and partial order follow the
// host builder's rules: touched rows of a block in first-touch order, then
// interior rows before shared ones, each group by decreasing slot count
// (ties by first touch); the slots of a row in element order; the partials
// of a shared row in block order.
constexpr int LB = 512;          // work-items of the layout kernels

__device__ inline uint32_t row_hash(int32_t r, int hs)
{
    return ((uint32_t)r * 2654435761u) & (uint32_t)(hs - 1);
}

// exclusive prefix sum of one int per work-item over the work-group (in
// work-item order); *total gets the sum
__device__ inline int group_scan_excl(int v, int *total)
{
    __shared__ int ws[LB / 64];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (l >= o)
            x += y;
    }
    if (l == 63)
        ws[w] = x;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < LB / 64; k++) {
        base += k < w ? ws[k] : 0;
        tot += ws[k];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

__global__ void k_lay_iperm(int64_t n, const int32_t *perm, int32_t *iperm)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        iperm[perm[i]] = (int32_t)i;
}

// element map in the solver numbering, elements in curve order, Dirichlet
// DOFs -1 (brow: Dirichlet flags in PHG's numbering)
__global__ void k_lay_emap(int64_t ne, int nb, const int32_t *eperm, const int32_t *map,
                           const int32_t *iperm, const uint8_t *brow, int32_t *emap)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne * nb)
        return;
    const int64_t e2 = t / nb, a = t % nb;
    const int32_t d = map[(int64_t)eperm[e2] * nb + a];
    emap[t] = brow[d] ? -1 : (iperm ? iperm[d] : d);
}

// per block: touched rows in first-touch order (rows_ft[b][k]), slots per
// touched row (ns_g[b][k]), the touched-row number of every entry (kft,
// KD_NONE for Dirichlet DOFs), blocks per row (cnt, atomically), rows per
// block (nk_g).  ovf is set when a block touches more than kmax rows.
template <int HS>
__global__ void __launch_bounds__(LB) k_lay_touch(int64_t ne, int nb, int eblk, int kmax,
                                                  const int32_t *__restrict__ emap, int32_t *cnt,
                                                  int32_t *rows_ft, int32_t *ns_g,
                                                  int32_t *nk_g, uint32_t *kft, int32_t *ovf)
{
    __shared__ int32_t hkey[HS], hmin[HS], hrank[HS];
    __shared__ int32_t nsc[HS / 2], rft[HS / 2];
    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x, e0 = b * eblk;
    const int nel = (int)min((int64_t)eblk, ne - e0), nt = nel * nb;
    const int32_t *em = emap + e0 * nb;
    for (int i = tid; i < HS; i += LB) {
        hkey[i] = -1;
        hmin[i] = 0x7fffffff;
        hrank[i] = -1;
    }
    for (int i = tid; i < HS / 2; i += LB)
        nsc[i] = 0;
    __syncthreads();
    // insert every row, keep its first position
    bool bad = false;
    for (int t = tid; t < nt && !bad; t += LB) {
        const int32_t r = em[t];
        if (r < 0)
            continue;
        uint32_t h = row_hash(r, HS);
        int probes = 0;
        while (true) {
            const int32_t k = hkey[h];
            if (k == r)
                break;
            if (k == -1) {
                const int32_t old = atomicCAS(&hkey[h], -1, r);
                if (old == -1 || old == r)
                    break;
            }
            h = (h + 1) & (HS - 1);
            if (++probes == HS) {
                bad = true;
                break;
            }
        }
        if (!bad)
            atomicMin(&hmin[h], t);
    }
    if (__syncthreads_or(bad)) {
        if (tid == 0)
            *ovf = 1;
        return;
    }
    auto find = [&](int32_t r) {
        uint32_t h = row_hash(r, HS);
        for (int probes = 0; hkey[h] != r && probes < HS; probes++)
            h = (h + 1) & (HS - 1);
        return h;
    };
    // first touches in position order: each work-item a contiguous range
    const int cw = (nt + LB - 1) / LB, t0 = min(nt, tid * cw), t1 = min(nt, t0 + cw);
    int mine = 0;
    for (int t = t0; t < t1; t++) {
        const int32_t r = em[t];
        if (r >= 0 && hmin[find(r)] == t)
            mine++;
    }
    int nk = 0;
    int k = group_scan_excl(mine, &nk);
    if (nk > kmax) {
        if (tid == 0)
            *ovf = 1;
        return;
    }
    for (int t = t0; t < t1; t++) {
        const int32_t r = em[t];
        if (r >= 0) {
            const uint32_t h = find(r);
            if (hmin[h] == t) {
                hrank[h] = k;
                rft[k] = r;
                k++;
            }
        }
    }
    __syncthreads();
    for (int t = tid; t < nt; t += LB) {
        const int32_t r = em[t];
        if (r < 0) {
            kft[e0 * nb + t] = KD_NONE;
            continue;
        }
        const int32_t kr = hrank[find(r)];
        kft[e0 * nb + t] = (uint32_t)kr;
        atomicAdd(&nsc[kr], 1);
    }
    __syncthreads();
    for (int i = tid; i < nk; i += LB) {
        rows_ft[b * kmax + i] = rft[i];
        ns_g[b * kmax + i] = nsc[i];
        atomicAdd(&cnt[rft[i]], 1);
    }
    if (tid == 0)
        nk_g[b] = nk;
}

// per block: the order of its touched rows (interior first, then shared,
// each by decreasing slot count, ties by first touch), rows, slot starts
// rs, interior count, the entries' row and slot (kd: row | slot << 16, the
// slots of a row in element order), and the sort keys of the partial order
// (shared rows: their row; interior: INT_MAX) with the entry index.
template <int KP>
__global__ void __launch_bounds__(LB) k_lay_order(int64_t ne, int nb, int eblk, int kmax,
                                                  const int32_t *__restrict__ cnt,
                                                  const int32_t *__restrict__ rows_ft,
                                                  const int32_t *__restrict__ ns_g,
                                                  const int32_t *__restrict__ nk_g,
                                                  const int32_t *__restrict__ eb_row,
                                                  uint32_t *kd, int32_t *rows, uint16_t *rs,
                                                  int32_t *eb_nint, int32_t *pkey,
                                                  int32_t *pval, int32_t *nint_tot)
{
    __shared__ uint64_t key[KP];
    __shared__ int32_t newpos[KP], sc[KP + 1], cmin[KP], ccnt[KP];
    constexpr int MAXE = 32;          // entries per work-item (nt <= LB * MAXE)
    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x, e0 = b * eblk;
    const int nel = (int)min((int64_t)eblk, ne - e0), nt = nel * nb;
    const int nk = nk_g[b];
    const int64_t k0 = eb_row[b];
    for (int k = tid; k < KP; k += LB) {
        uint64_t v = ~0ull;
        if (k < nk) {
            const int32_t r = rows_ft[b * kmax + k];
            const uint64_t shared = cnt[r] > 1 ? 1 : 0;
            v = shared << 40 | (uint64_t)(0xfffff - ns_g[b * kmax + k]) << 20 | (uint64_t)k;
        }
        key[k] = v;
        cmin[k] = 0x7fffffff;
        ccnt[k] = 0;
    }
    __syncthreads();
    // bitonic sort of the KP keys (unique below the padding)
    for (int sz = 2; sz <= KP; sz <<= 1)
        for (int st = sz >> 1; st > 0; st >>= 1) {
            for (int i = tid; i < KP / 2; i += LB) {
                const int lo = 2 * i - (i & (st - 1)), hi = lo + st;
                const bool up = (lo & sz) == 0;
                const uint64_t a = key[lo], c = key[hi];
                if ((a > c) == up) {
                    key[lo] = c;
                    key[hi] = a;
                }
            }
            __syncthreads();
        }
    // sorted position j of first-touch row k; slots per sorted row
    const int per = KP / LB;          // consecutive sorted rows per work-item
    int mine = 0, ni = 0;
    for (int j = tid * per; j < tid * per + per; j++)
        if (j < nk) {
            const uint64_t v = key[j];
            const int k = (int)(v & 0xfffff);
            newpos[k] = j;
            const int32_t r = rows_ft[b * kmax + k];
            rows[k0 + j] = r;
            const bool shared = (v >> 40) != 0;
            pkey[k0 + j] = shared ? r : 0x7fffffff;
            pval[k0 + j] = (int32_t)(k0 + j);
            ni += shared ? 0 : 1;
            mine += 0xfffff - (int)((v >> 20) & 0xfffff);
        }
    int tot = 0;
    int s = group_scan_excl(mine, &tot);
    for (int j = tid * per; j < tid * per + per; j++)
        if (j < nk) {
            sc[j] = s;
            rs[k0 + b + j] = (uint16_t)s;
            s += 0xfffff - (int)((key[j] >> 20) & 0xfffff);
        }
    int nitot = 0;
    (void)group_scan_excl(ni, &nitot);
    if (tid == 0) {
        rs[k0 + b + nk] = (uint16_t)tot;
        eb_nint[b] = nitot;
        atomicAdd(nint_tot, nitot);
    }
    __syncthreads();
    // slots of each row in element order: round after round, the first
    // entry still waiting takes the row's next slot
    int kl[MAXE];
    uint32_t act = 0;
#pragma unroll
    for (int i = 0; i < MAXE; i++) {
        const int t = tid + LB * i;
        kl[i] = -1;
        if (t < nt) {
            const uint32_t f = kd[e0 * nb + t];
            if (f == KD_NONE)
                continue;
            kl[i] = newpos[f];
            act |= 1u << i;
        }
    }
    while (__syncthreads_or(act != 0)) {
#pragma unroll
        for (int i = 0; i < MAXE; i++)
            if (act >> i & 1u)
                atomicMin(&cmin[kl[i]], tid + LB * i);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < MAXE; i++) {
            const int t = tid + LB * i;
            if ((act >> i & 1u) && cmin[kl[i]] == t) {
                kd[e0 * nb + t] = (uint32_t)kl[i] | (uint32_t)(sc[kl[i]] + ccnt[kl[i]]) << 16;
                ccnt[kl[i]]++;
                cmin[kl[i]] = 0x7fffffff;
                act &= ~(1u << i);
            }
        }
    }
}

// partial positions: pkey / pval sorted by row (stably, so in block order);
// the first npart are the shared entries.  Starts of rows give fx_rows and
// fx_off (through the scan of flag), ppos[entry] = its sorted position.
__global__ void k_lay_flags(int64_t npart, const int32_t *pkey, int32_t *flag)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < npart)
        flag[j] = (j == 0 || pkey[j] != pkey[j - 1]) ? 1 : 0;
    else if (j == npart)
        flag[j] = 0;
}

__global__ void k_lay_partials(int64_t npart, const int32_t *pkey, const int32_t *pval,
                               const int32_t *flag, const int32_t *fpos, int32_t *ppos,
                               int32_t *fx_rows, int32_t *fx_off, int32_t *maxcnt)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npart)
        return;
    ppos[pval[j]] = (int32_t)j;
    if (flag[j]) {
        const int32_t u = fpos[j];
        fx_rows[u] = pkey[j];
        fx_off[u] = (int32_t)j;
        int64_t e = j + 1;
        while (e < npart && !flag[e])
            e++;
        atomicMax(maxcnt, (int32_t)(e - j));
    }
}

// Dirichlet rows in the solver numbering, in row order
__global__ void k_lay_dflags(int64_t n, const int32_t *perm, const uint8_t *brow, uint8_t *f)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        f[i] = brow[perm ? perm[i] : i] ? 1 : 0;
}

// lane-group index chunks (see load_round), from kd
__global__ void k_lay_chunks(int64_t ne, int p, int nb, const uint32_t *kd, uint32_t *idx)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne * 4)
        return;
    const int64_t e = t / 4;
    const int g = (int)(t % 4);
    const int nj = (nb + 3) / 4, no = (nb / 2 + 1) / 2, wc = (nj + no + 1) / 2;
    const int rec = phg::eop_record_dwords(p);
    uint32_t h[10];
    for (int u = 0; u < 10; u++)
        h[u] = 0xffffu;
    for (int u = 0; u < nj; u++) {
        const int j = p == 3 ? nj * g + u : 4 * u + g;
        if (j < nb)
            h[u] = kd[e * nb + j] & 0xffffu;
    }
    for (int o = 0; o < no; o++) {
        int i = -1;
        if (p == 2)
            i = kOutRow2(4 * g + o);
        else {
            const int tt = 2 * o + (g & 1);
            if (tt < nb / 2)
                i = 2 * tt + (g >> 1);
        }
        if (i >= 0)
            h[nj + o] = kd[e * nb + i] >> 16;
    }
    if (p == 2 && g >= 2) {
        uint32_t *d = idx + e * rec + 2 * g + 2;
        d[0] = h[0] | h[1] << 16;
        d[1] = h[nj] | h[nj + 1] << 16;
        return;
    }
    const int off = p == 2 ? 3 * g : g * wc;
    for (int u = 0; u < wc; u++)
        idx[e * rec + off + u] = h[2 * u] | h[2 * u + 1] << 16;
}

// per-row codes of the shared rows: partial start | count << 24
__global__ void k_lay_codes(int64_t nfx, const int32_t *fx_rows, const int32_t *fx_off,
                            uint32_t *code)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nfx)
        code[fx_rows[u]] = (uint32_t)fx_off[u] | (uint32_t)(fx_off[u + 1] - fx_off[u]) << 24;
}

__global__ void k_lay_fill_u32(int64_t n, uint32_t v, uint32_t *a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = v;
}

// device temporaries of the layout build, freed on every exit
struct DevTemps {
    std::vector<void *> p;
    template <typename T>
    int get(T **d, size_t count)
    {
        *d = nullptr;
        const int rc = phg::dmalloc((void **)d, std::max<size_t>(count * sizeof(T), 4));
        if (rc == 0)
            p.push_back((void *)*d);
        return rc;
    }
    ~DevTemps()
    {
        for (void *q : p)
            phg::dfree(q);
    }
};

template <typename T>
int read_one(T *h, const T *d)
{
    return phg::download(h, d, sizeof(T));
}

// The layout of the blocks of EBLK consecutive elements in curve order
// (d_eperm: that order, taken over by the layout), built on the device.
// Returns PHG_HIP_OK, 1 when some block touches more than KMAX rows (the
// host's split builds the layout then; nothing is kept), or an error.
template <int P, typename Mark>
int build_layout_device(EopLayout &L, const phg_hip_space *sp, const phg_hip_mat *A,
                        int32_t *d_eperm, int codes, hipStream_t s, Mark &&mark)
{
    using S = EopShape<P>;
    constexpr int NB = S::NB, EBLK = S::EBLK, KMAX = S::KMAX;
    constexpr int HS = P == 1 ? 2048 : (P == 2 ? 4096 : 8192);
    constexpr int KP = P == 1 ? 1024 : (P == 2 ? 2048 : 4096);
    static_assert(2 * KMAX <= HS && KMAX <= KP && NB * EBLK <= 32 * LB, "layout kernel sizes");
    using namespace phg;
    int rc;
    const int64_t ne = sp->nelem, n = A->n;
    const int64_t nblk = (ne + EBLK - 1) / EBLK;
    const int nbk = (int)nblk;
    auto grid = [](int64_t m) { return dim3((unsigned)std::max<int64_t>(1, (m + 255) / 256)); };
    DevTemps tmp;
    int32_t *iperm = nullptr, *emap = nullptr, *cnt = nullptr, *rows_ft = nullptr,
            *ns_g = nullptr, *nk_g = nullptr, *stats = nullptr;
    uint32_t *kd = nullptr;
    if ((A->q_perm && (rc = tmp.get(&iperm, (size_t)n))) ||
        (rc = tmp.get(&emap, (size_t)(ne * NB))) || (rc = tmp.get(&cnt, (size_t)n)) ||
        (rc = tmp.get(&rows_ft, (size_t)(nblk * KMAX))) ||
        (rc = tmp.get(&ns_g, (size_t)(nblk * KMAX))) ||
        (rc = tmp.get(&nk_g, (size_t)(nblk + 1))) || (rc = tmp.get(&stats, 8)) ||
        (rc = tmp.get(&kd, (size_t)(ne * NB))))
        return rc;
    PHG_HIP_CHECK(hipMemsetAsync(cnt, 0, (size_t)n * 4, s));
    PHG_HIP_CHECK(hipMemsetAsync(nk_g, 0, (size_t)(nblk + 1) * 4, s));
    PHG_HIP_CHECK(hipMemsetAsync(stats, 0, 8 * 4, s));
    if (iperm)
        hipLaunchKernelGGL(k_lay_iperm, grid(n), dim3(256), 0, s, n, (const int32_t *)A->q_perm,
                           iperm);
    hipLaunchKernelGGL(k_lay_emap, grid(ne * NB), dim3(256), 0, s, ne, NB,
                       (const int32_t *)d_eperm, (const int32_t *)sp->d_map,
                       (const int32_t *)iperm, (const uint8_t *)A->d_bdry_row, emap);
    hipLaunchKernelGGL((k_lay_touch<HS>), dim3((unsigned)nblk), dim3(LB), 0, s, ne, NB, EBLK,
                       KMAX, (const int32_t *)emap, cnt, rows_ft, ns_g, nk_g, kd, stats);
    PHG_HIP_CHECK(hipGetLastError());
    int32_t ovf = 0;
    if ((rc = read_one(&ovf, stats)))
        return rc;
    mark("layout: element map, touched rows");
    if (ovf)
        return 1;
    // block row ranges
    void *cub = nullptr;
    size_t cub_bytes = 0, need = 0;
    auto cub_reserve = [&](size_t bytes) -> int {
        if (bytes <= cub_bytes)
            return PHG_HIP_OK;
        cub_bytes = bytes;
        return tmp.get((uint8_t **)&cub, bytes);
    };
    PHG_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, nk_g, nk_g, nbk + 1, s));
    if ((rc = cub_reserve(need)) || (rc = dmalloc((void **)&L.eb_row, (size_t)(nblk + 1) * 4)))
        return rc;
    PHG_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cub, need, nk_g, L.eb_row, nbk + 1, s));
    int32_t nent32 = 0;
    if ((rc = read_one(&nent32, L.eb_row + nblk)))
        return rc;
    const int64_t nent = nent32;
    int32_t *pkey = nullptr, *pval = nullptr, *pkey2 = nullptr, *pval2 = nullptr;
    if ((rc = dmalloc((void **)&L.rows, (size_t)(nent + 1) * 4)) ||
        (rc = dmalloc((void **)&L.rs, (size_t)(nent + nblk + 1) * 2)) ||
        (rc = dmalloc((void **)&L.eb_nint, (size_t)nblk * 4)) ||
        (rc = dmalloc((void **)&L.ppos, (size_t)(nent + 1) * 4)) ||
        (rc = tmp.get(&pkey, (size_t)nent + 1)) || (rc = tmp.get(&pval, (size_t)nent + 1)) ||
        (rc = tmp.get(&pkey2, (size_t)nent + 1)) || (rc = tmp.get(&pval2, (size_t)nent)))
        return rc;
    hipLaunchKernelGGL((k_lay_order<KP>), dim3((unsigned)nblk), dim3(LB), 0, s, ne, NB, EBLK,
                       KMAX, (const int32_t *)cnt, (const int32_t *)rows_ft,
                       (const int32_t *)ns_g, (const int32_t *)nk_g,
                       (const int32_t *)L.eb_row, kd, L.rows, L.rs, L.eb_nint, pkey, pval,
                       stats + 1);
    PHG_HIP_CHECK(hipGetLastError());
    // partials of the shared rows in block order
    PHG_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, pkey, pkey2, pval, pval2,
                                                     (int)nent, 0, 32, s));
    if ((rc = cub_reserve(need)))
        return rc;
    PHG_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub, need, pkey, pkey2, pval, pval2,
                                                     (int)nent, 0, 32, s));
    int32_t nint32 = 0;
    if ((rc = read_one(&nint32, stats + 1)))
        return rc;
    mark("layout: row order, slots, sort");
    const int64_t nint = nint32, npart = nent - nint;
    int32_t *flag = pkey, *fpos = pval;     // the unsorted pairs are spent
    hipLaunchKernelGGL(k_lay_flags, grid(npart + 1), dim3(256), 0, s, npart,
                       (const int32_t *)pkey2, flag);
    PHG_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, flag, fpos, (int)npart + 1, s));
    if ((rc = cub_reserve(need)))
        return rc;
    PHG_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cub, need, flag, fpos, (int)npart + 1, s));
    int32_t nfx32 = 0;
    if ((rc = read_one(&nfx32, fpos + npart)))
        return rc;
    const int64_t nfx = nfx32;
    // Dirichlet rows after the shared rows, in row order
    uint8_t *dflag = nullptr;
    int32_t *dsel = nullptr;
    if ((rc = tmp.get(&dflag, (size_t)n)) || (rc = tmp.get(&dsel, (size_t)n)))
        return rc;
    hipLaunchKernelGGL(k_lay_dflags, grid(n), dim3(256), 0, s, n, (const int32_t *)A->q_perm,
                       (const uint8_t *)A->d_bdry_row, dflag);
    hipcub::CountingInputIterator<int32_t> ci(0);
    PHG_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, need, ci, dflag, dsel, stats + 2,
                                                (int)n, s));
    if ((rc = cub_reserve(need)))
        return rc;
    PHG_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub, need, ci, dflag, dsel, stats + 2, (int)n,
                                                s));
    int32_t ndir32 = 0;
    if ((rc = read_one(&ndir32, stats + 2)))
        return rc;
    const int64_t ndir = ndir32;
    if ((rc = dmalloc((void **)&L.fx_rows, (size_t)(nfx + ndir) * 4)) ||
        (rc = dmalloc((void **)&L.fx_off, (size_t)(nfx + 1) * 4)))
        return rc;
    hipLaunchKernelGGL(k_lay_fill_u32, grid(nent + 1), dim3(256), 0, s, nent + 1, 0xffffffffu,
                       (uint32_t *)L.ppos);
    hipLaunchKernelGGL(k_lay_partials, grid(npart), dim3(256), 0, s, npart,
                       (const int32_t *)pkey2, (const int32_t *)pval2, (const int32_t *)flag,
                       (const int32_t *)fpos, L.ppos, L.fx_rows, L.fx_off, stats + 3);
    PHG_HIP_CHECK(hipGetLastError());
    const int32_t np32 = (int32_t)npart, zero = 0;
    if ((rc = upload(L.fx_off + nfx, &np32, 4)) || (rc = upload(L.rows + nent, &zero, 4)) ||
        (rc = upload(L.ppos + nent, &zero, 4)))
        return rc;
    PHG_HIP_CHECK(hipMemsetAsync(L.rs + nent + nblk, 0, 2, s));
    if (ndir > 0)
        PHG_HIP_CHECK(hipMemcpyAsync(L.fx_rows + nfx, dsel, (size_t)ndir * 4,
                                     hipMemcpyDeviceToDevice, s));
    // element ranges, index chunks
    std::vector<int32_t> eb_elem((size_t)nblk + 1);
    for (int64_t b = 0; b <= nblk; b++)
        eb_elem[(size_t)b] = (int32_t)std::min<int64_t>(b * EBLK, ne);
    const int rec = eop_record_dwords(P);
    if ((rc = dmalloc((void **)&L.eb_elem, (size_t)(nblk + 1) * 4)) ||
        (rc = upload(L.eb_elem, eb_elem.data(), (size_t)(nblk + 1) * 4)) ||
        (rc = dmalloc((void **)&L.idx, (size_t)ne * rec * 4)))
        return rc;
    hipLaunchKernelGGL(k_lay_chunks, grid(ne * 4), dim3(256), 0, s, ne, P, NB,
                       (const uint32_t *)kd, L.idx);
    PHG_HIP_CHECK(hipGetLastError());
    mark("layout: partials, Dirichlet rows");
    // per-row codes when every partial start and count fits
    int32_t maxcnt = 0;
    if ((rc = read_one(&maxcnt, stats + 3)))
        return rc;
    if (codes && npart < ((int64_t)1 << 24) - 1 && maxcnt < 256) {
        if ((rc = dmalloc((void **)&L.fxc, (size_t)n * 4)))
            return rc;
        hipLaunchKernelGGL(k_lay_fill_u32, grid(n), dim3(256), 0, s, n, EOP_NOT_SHARED, L.fxc);
        hipLaunchKernelGGL(k_lay_codes, grid(nfx), dim3(256), 0, s, nfx,
                           (const int32_t *)L.fx_rows, (const int32_t *)L.fx_off, L.fxc);
    }
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    mark("layout: chunks, codes");
    L.nelem = ne;
    L.n = n;
    L.nblk = nblk;
    L.nent = nent;
    L.nint = nint;
    L.nfx = nfx;
    L.ndir = ndir;
    L.npart = npart;
    L.eperm = d_eperm;            // the layout holds the order from here on
    return PHG_HIP_OK;
}

}  // namespace

//============================================================== C ABI ====

extern "C" int phg_hip_mat_element_operator(phg_hip_mat *A,
                                            const phg_hip_space *sp,
                                            const phg_hip_problem *pb,
                                            const double *kappa)
{
    using namespace phg;
    int rc;
    if ((rc = require_ctx()))
        return rc;
    if (A == nullptr || sp == nullptr || pb == nullptr)
        return set_error(PHG_HIP_E_ARG, "phg_hip_mat_element_operator: null argument");
    if (!A->eliminated || A->dim != 1 || pb->dim != 1 || A->n != sp->ndof ||
        sp->order < 1 || sp->order > 3 || (pb->use_kappa && kappa == nullptr) ||
        !std::isfinite(pb->mass_coef))
        return set_error(PHG_HIP_E_ARG,
                         "phg_hip_mat_element_operator: needs the eliminated matrix "
                         "of a scalar problem (stiffness, optional mass term) on this space");
    // PHG_HIP_SETUP_TRACE=1 prints the time of each set-up phase to stderr
    static const bool trace = getenv("PHG_HIP_SETUP_TRACE") != nullptr;
    auto t_last = std::chrono::steady_clock::now();
    auto mark = [&](const char *what) {
        if (!trace)
            return;
        const auto t = std::chrono::steady_clock::now();
        fprintf(stderr, "element operator set-up: %-22s %8.2f ms\n", what,
                std::chrono::duration<double, std::milli>(t - t_last).count());
        t_last = t;
    };
    if (A->batch_refs > 0)
        return set_error(PHG_HIP_E_STATE, "phg_hip_mat_element_operator: the matrix is in a "
                                          "batch, whose graphs hold its operator");
    if ((rc = ensure_solver_copy(A)) || (rc = ensure_fp32_copy(A)))
        return rc;
    free_eop(A);
    mark("fp32 copy");
    hipStream_t s = ctx()->stream;
    const int p = sp->order, nb = sp->nbas;
    const int64_t ne = sp->nelem, n = A->n;
    if ((int64_t)nb * ne >= (int64_t)1 << 30)
        return set_error(PHG_HIP_E_ARG, "phg_hip_mat_element_operator: too many elements");
    // per-row codes of the shared rows for the solvers' update kernel, unless
    // PHG_HIP_EOP_NO_ROW_CODES is set (then the fix-up kernel sums them)
    const int codes = getenv("PHG_HIP_EOP_NO_ROW_CODES") == nullptr ? 1 : 0;
    // A matrix assembled on this space has the space's Dirichlet rows and
    // solver numbering (both functions of the mesh alone), so it shares the
    // space's layout; only the coefficients depend on kappa.
    const bool same_space = A->space_uid != 0 && A->space_uid == sp->uid;
    auto build_layout = [&](EopLayout &L) -> int {
    auto put = [&](void **d, const void *h, size_t bytes) {
        int r2 = dmalloc(d, std::max<size_t>(bytes, 4));
        if (r2 == 0 && bytes > 0)
            r2 = upload(*d, h, bytes);
        return r2;
    };
    // both builders: A fragments, mass row sums, launch grids
    auto finish = [&]() -> int {
    int ks = p == 1 ? EopShape<1>::KS : (p == 2 ? EopShape<2>::KS : EopShape<3>::KS);
    // A fragments: bf16 parts (P3) or FP32 values (P1, P2), as bytes;
    // without and with the mass tensor
    std::vector<uint8_t> tiles[2];
    for (int ms = 0; ms < 2; ms++) {
        if (p == 3) {
            std::vector<uint16_t> tb;
            build_tiles(p, nb, ks, ms != 0, tb);
            tiles[ms].resize(tb.size() * 2);
            memcpy(tiles[ms].data(), tb.data(), tiles[ms].size());
        }
        else if (p == 2) {
            std::vector<float> tf;
            if (!build_tiles_pts(p, nb, tf, L.msum))
                return set_error(PHG_HIP_E_STATE,
                                 "element operator set-up: the P2 rule has not %d points", kNp2);
            tiles[ms].resize(tf.size() * 4);
            memcpy(tiles[ms].data(), tf.data(), tiles[ms].size());
        }
        else {
            std::vector<float> tf;
            build_tiles_f32(p, nb, ms != 0, tf);
            tiles[ms].resize(tf.size() * 4);
            memcpy(tiles[ms].data(), tf.data(), tiles[ms].size());
        }
    }
    // row sums of the mass tensor; the local indices 2 t, 2 t + 1 of a tile
    // are of one kind (vertex, edge or face nodes: the kinds start at even
    // local indices), so their sums agree and one value serves the tile
    // (P2: per output row, set with the tiles)
    if (p != 2) {
        std::vector<double> T((size_t)NREF * nb * nb);
        eop_reference_tensors(p, T.data());
        for (int t = 0; t < nb / 2; t++) {
            double r[2] = {0.0, 0.0};
            for (int h = 0; h < 2; h++)
                for (int j = 0; j < nb; j++)
                    r[h] += T[((size_t)6 * nb + 2 * t + h) * nb + j];
            if (std::fabs(r[0] - r[1]) > 1e-14)
                return set_error(PHG_HIP_E_STATE,
                                 "element operator set-up: mass row sums of tile %d differ", t);
            L.msum[t] = (float)r[0];
        }
    }
    L.p = p;
    L.nb = nb;
    L.ks = ks;
    if ((rc = put((void **)&L.tiles, tiles[0].data(), tiles[0].size())) ||
        (rc = put((void **)&L.tiles_mass, tiles[1].data(), tiles[1].size())))
        return rc;
    // grids: every block group resident; partial sums of both kernels fit
    // the solvers' partial-sum segment
    const int gb = p == 1 ? grid_block<1, false>()
                          : (p == 2 ? grid_block<2, false>() : grid_block<3, false>());
    L.grid_b = (int)std::max<int64_t>(1, std::min<int64_t>({L.nblk, (int64_t)gb,
                                                             (int64_t)EOP_MAX_PART}));
    L.grid_f = (int)std::max<int64_t>(1, std::min<int64_t>((L.nfx + FB - 1) / FB, 4096));
    mark("tiles, grids");
    return PHG_HIP_OK;
    };
    // elements in Hilbert-curve order of their centroids
    std::vector<int32_t> eperm((size_t)ne);
    bool on_device = false;
    {
        // the space's bounding box of its vertices
        const double *lo = sp->lo, *hi = sp->hi;
        double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        const double sc = ext > 0 ? 2097151.0 / ext : 0.0;
        // (key, element) pairs radix-sorted on the device: a stable sort by
        // key, as the elements start in index order
        uint64_t *dk = nullptr, *dk2 = nullptr;
        int32_t *di = nullptr, *di2 = nullptr;
        void *tmp = nullptr;
        size_t tmp_bytes = 0;
        auto done_sort = [&]() {
            dfree(dk); dfree(dk2); dfree(di); dfree(di2); dfree(tmp);
        };
        if ((rc = dmalloc((void **)&dk, (size_t)ne * 8)) ||
            (rc = dmalloc((void **)&dk2, (size_t)ne * 8)) ||
            (rc = dmalloc((void **)&di, (size_t)ne * 4)) ||
            (rc = dmalloc((void **)&di2, (size_t)ne * 4))) {
            done_sort();
            return rc;
        }
        hipLaunchKernelGGL(k_elem_morton, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s,
                           ne, (const double *)sp->d_verts, (const int32_t *)sp->d_elem_verts,
                           lo[0], lo[1], lo[2], sc, dk, di);
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, dk, dk2, di, di2, (int)ne,
                                               0, 63, s) != hipSuccess ||
            (rc = dmalloc(&tmp, tmp_bytes)) ||
            hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, dk, dk2, di, di2, (int)ne, 0,
                                               63, s) != hipSuccess) {
            done_sort();
            return rc ? rc : set_error(PHG_HIP_E_HIP, "element order: radix sort failed");
        }
        mark("curve keys, sort");
        // blocks of EBLK consecutive elements built on the device when each
        // touches at most KMAX rows (PHG_HIP_EOP_HOST_LAYOUT: the host split)
        if (getenv("PHG_HIP_EOP_HOST_LAYOUT") == nullptr) {
            rc = p == 1 ? build_layout_device<1>(L, sp, A, di2, codes, s, mark)
                        : (p == 2 ? build_layout_device<2>(L, sp, A, di2, codes, s, mark)
                                  : build_layout_device<3>(L, sp, A, di2, codes, s, mark));
            if (rc == PHG_HIP_OK) {
                di2 = nullptr;        // the layout holds the order
                on_device = true;
            }
            else if (rc != 1) {
                done_sort();
                return rc;
            }
            rc = PHG_HIP_OK;
        }
        if (!on_device)
            rc = download(eperm.data(), di2, (size_t)ne * 4);
        done_sort();
        if (rc)
            return rc;
    }
    if (on_device) {
        L.on_device = 1;
        mark("device layout");
        return finish();
    }
    mark("element order download");
    std::vector<int32_t> perm;
    std::vector<uint8_t> brow((size_t)n);
    if ((rc = download(brow.data(), A->d_bdry_row, (size_t)n)))
        return rc;
    // solver numbering: solver index -> PHG index (identity without q_perm)
    std::vector<int32_t> iperm((size_t)n);
    if (A->q_perm) {
        perm.resize((size_t)n);
        if ((rc = download(perm.data(), A->q_perm, (size_t)n * 4)))
            return rc;
        for (int64_t i = 0; i < n; i++)
            iperm[(size_t)perm[(size_t)i]] = (int32_t)i;
    }
    else
        std::iota(iperm.begin(), iperm.end(), 0);
    // the element maps (PHG numbering), for the host split only
    HostVec<int32_t> map((size_t)ne * nb);
    if ((rc = download(map.data(), sp->d_map, (size_t)ne * nb * 4)))
        return rc;
    mark("downloads");
    // element DOF map in the solver numbering, Dirichlet DOFs -1
    HostVec<int32_t> emap((size_t)ne * nb);
    std::vector<uint8_t> bq((size_t)n);
    for (int64_t i = 0; i < n; i++)
        bq[(size_t)i] = brow[(size_t)(A->q_perm ? perm[(size_t)i] : i)];
    host_parallel(ne, [&](int64_t lo, int64_t hi, int) {
        for (int64_t e2 = lo; e2 < hi; e2++) {
            const int64_t e = eperm[(size_t)e2];
            for (int a = 0; a < nb; a++) {
                const int32_t r = iperm[(size_t)map[(size_t)(e * nb + a)]];
                emap[(size_t)(e2 * nb + a)] = bq[(size_t)r] ? -1 : r;
            }
        }
    });
    mark("element map");
    // element blocks: consecutive elements, at most eblk of them touching
    // at most kmax rows; the touched rows of a block in first-touch order.
    // The greedy split runs on its own in each of nch fixed element ranges
    // (one per 65,536 elements, at most 16), so the layout does not depend
    // on the number of host threads.
    const int eblk = p == 1 ? EopShape<1>::EBLK : (p == 2 ? EopShape<2>::EBLK : EopShape<3>::EBLK);
    const int kmax = p == 1 ? EopShape<1>::KMAX : (p == 2 ? EopShape<2>::KMAX : EopShape<3>::KMAX);
    const int64_t nch = std::max<int64_t>(1, std::min<int64_t>(16, ne / 65536));
    struct Split {
        std::vector<int32_t> start, nrow, rows;   // per block: first element, rows
    };
    std::vector<Split> split((size_t)nch);
    host_parallel(nch, [&](int64_t c0, int64_t c1, int) {
        std::unique_ptr<int32_t[]> stamp(new int32_t[(size_t)n]);
        std::fill(stamp.get(), stamp.get() + n, -1);
        int32_t b = -1;   // block tag in stamp, unique within this thread
        for (int64_t c = c0; c < c1; c++) {
            Split &sl = split[(size_t)c];
            int32_t nk = 0, nel = 0;
            for (int64_t e = ne * c / nch; e < ne * (c + 1) / nch; e++) {
                const int32_t *me = &emap[(size_t)(e * nb)];
                int fresh = 0;
                for (int a = 0; a < nb; a++)
                    if (me[a] >= 0 && stamp[(size_t)me[a]] != b)
                        fresh++;
                if (nel == 0 || nel == eblk || nk + fresh > kmax) {
                    if (nel > 0)
                        sl.nrow.push_back(nk);
                    sl.start.push_back((int32_t)e);
                    b++;
                    nk = nel = 0;
                }
                for (int a = 0; a < nb; a++) {
                    const int32_t r = me[a];
                    if (r >= 0 && stamp[(size_t)r] != b) {
                        stamp[(size_t)r] = b;
                        sl.rows.push_back(r);
                        nk++;
                    }
                }
                nel++;
            }
            if (nel > 0)
                sl.nrow.push_back(nk);
        }
    }, 1);
    std::vector<int32_t> eb_elem, eb_row{0}, rows;
    for (const Split &sl : split) {
        for (size_t t = 0; t < sl.start.size(); t++) {
            eb_elem.push_back(sl.start[t]);
            eb_row.push_back(eb_row.back() + sl.nrow[t]);
        }
        rows.insert(rows.end(), sl.rows.begin(), sl.rows.end());
    }
    eb_elem.push_back((int32_t)ne);
    // number of blocks touching each row
    std::vector<int32_t> cnt((size_t)n, 0);
    for (const int32_t r : rows)
        cnt[(size_t)r]++;
    const int64_t nblk = (int64_t)eb_elem.size() - 1, nent = (int64_t)rows.size();
    if (nblk <= 0)
        return set_error(PHG_HIP_E_ARG, "phg_hip_mat_element_operator: no elements");
    // per block: interior rows (touched by this block only) first, then
    // shared rows; slots of a row consecutive, in element order; the
    // partials of a shared row in block order
    std::vector<int32_t> eb_nint((size_t)nblk), ppos((size_t)nent, -1);
    std::vector<uint16_t> rs((size_t)(nent + nblk));
    HostVec<uint32_t> kd((size_t)ne * nb);   // every entry set per block
    std::vector<int32_t> fxi((size_t)n, -1), seen((size_t)n, 0), fx_rows, fx_off{0};
    // shared rows in row order (the fix-up walks them so), each with the
    // range of its partials
    for (int64_t r = 0; r < n; r++)
        if (cnt[(size_t)r] > 1) {
            fxi[(size_t)r] = (int32_t)fx_rows.size();
            fx_rows.push_back((int32_t)r);
            fx_off.push_back(fx_off.back() + cnt[(size_t)r]);
        }
    mark("blocks");
    // blocks are independent here (disjoint ranges of rows, rs and kd); each
    // host thread keeps its own row -> local index map
    host_parallel(nblk, [&](int64_t b0, int64_t b1, int) {
        std::unique_ptr<int32_t[]> kid(new int32_t[(size_t)n]);
        std::vector<int32_t> first, ns, ord, sc;
        for (int64_t b = b0; b < b1; b++) {
            const int32_t k0 = eb_row[(size_t)b], nk = eb_row[(size_t)b + 1] - k0;
            const int32_t e0 = eb_elem[(size_t)b], e1 = eb_elem[(size_t)b + 1];
            // slots per touched row (first-touch order)
            first.assign(rows.begin() + k0, rows.begin() + k0 + nk);
            ns.assign((size_t)nk, 0);
            for (int32_t k = 0; k < nk; k++)
                kid[(size_t)first[(size_t)k]] = k;
            for (int64_t e = e0; e < e1; e++)
                for (int a = 0; a < nb; a++) {
                    const int32_t r = emap[(size_t)(e * nb + a)];
                    if (r >= 0)
                        ns[(size_t)kid[(size_t)r]]++;
                }
            // interior rows first, then shared; each group by decreasing slot
            // count, so the row sums of a wavefront have similar lengths
            ord.resize((size_t)nk);
            std::iota(ord.begin(), ord.end(), 0);
            std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t c) {
                const bool ia = cnt[(size_t)first[(size_t)a]] == 1,
                           ic = cnt[(size_t)first[(size_t)c]] == 1;
                if (ia != ic)
                    return ia;
                return ns[(size_t)a] > ns[(size_t)c];
            });
            int32_t ni = 0;
            for (int32_t k = 0; k < nk; k++) {
                const int32_t r = first[(size_t)ord[(size_t)k]];
                rows[(size_t)(k0 + k)] = r;
                if (cnt[(size_t)r] == 1)
                    ni++;
            }
            eb_nint[(size_t)b] = ni;
            int32_t k;
            for (k = 0; k < nk; k++)
                kid[(size_t)rows[(size_t)(k0 + k)]] = k;
            sc.assign((size_t)nk + 1, 0);
            for (int64_t e = e0; e < e1; e++)
                for (int a = 0; a < nb; a++) {
                    const int32_t r = emap[(size_t)(e * nb + a)];
                    if (r >= 0)
                        sc[(size_t)kid[(size_t)r] + 1]++;
                }
            for (k = 0; k < nk; k++)
                sc[(size_t)k + 1] += sc[(size_t)k];
            for (k = 0; k <= nk; k++)
                rs[(size_t)(k0 + b + k)] = (uint16_t)sc[(size_t)k];
            for (int64_t e = e0; e < e1; e++)
                for (int a = 0; a < nb; a++) {
                    const int32_t r = emap[(size_t)(e * nb + a)];
                    if (r >= 0) {
                        const int32_t kl = kid[(size_t)r];
                        kd[(size_t)(e * nb + a)] = (uint32_t)kl | (uint32_t)sc[(size_t)kl]++ << 16;
                    }
                    else
                        kd[(size_t)(e * nb + a)] = KD_NONE;
                }
        }
    }, 16);
    // partials of the shared rows in block order
    int64_t nint = 0;
    for (int64_t b = 0; b < nblk; b++) {
        const int32_t k0 = eb_row[(size_t)b], nk = eb_row[(size_t)b + 1] - k0;
        nint += eb_nint[(size_t)b];
        for (int32_t k = eb_nint[(size_t)b]; k < nk; k++) {
            const int32_t r = rows[(size_t)(k0 + k)];
            ppos[(size_t)(k0 + k)] = fx_off[(size_t)fxi[(size_t)r]] + seen[(size_t)r]++;
        }
    }
    mark("block tables");
    // lane-group index chunks (see load_round): x rows of j = NJ g + u
    // (P3) or 4 u + g, result slots of i = 2 t + g / 2, t = 2 o + g % 2 (P1,
    // P3) or kOutRow2(4 g + o) (P2)
    const int nj = (nb + 3) / 4, no = (nb / 2 + 1) / 2, wc = (nj + no + 1) / 2;
    const int rec = eop_record_dwords(p);
    HostVec<uint32_t> idx((size_t)ne * rec);
    host_parallel(ne, [&](int64_t lo, int64_t hi, int) {
        for (int64_t e = lo; e < hi; e++)
            for (int g = 0; g < 4; g++) {
                uint32_t h[16];   // 2 wc <= 10 halves (P3: nj = no = 5, wc = 5)
                for (int u = 0; u < 2 * wc; u++)
                    h[u] = 0xffffu;
                for (int u = 0; u < nj; u++) {
                    const int j = p == 3 ? nj * g + u : 4 * u + g;
                    if (j < nb)
                        h[u] = kd[(size_t)(e * nb + j)] & 0xffffu;
                }
                for (int o = 0; o < no; o++) {
                    if (p == 2) {
                        const int i = kOutRow2(4 * g + o);
                        if (i >= 0)
                            h[nj + o] = kd[(size_t)(e * nb + i)] >> 16;
                        continue;
                    }
                    const int t = 2 * o + (g & 1);
                    if (t < nb / 2)
                        h[nj + o] = kd[(size_t)(e * nb + 2 * t + (g >> 1))] >> 16;
                }
                if (p == 2 && g >= 2) {
                    // j 4 u + g is past nb for u = 2, i past the lane's two
                    uint32_t *d = &idx[(size_t)(e * rec + 2 * g + 2)];
                    d[0] = h[0] | h[1] << 16;
                    d[1] = h[nj] | h[nj + 1] << 16;
                    continue;
                }
                const int off = p == 2 ? 3 * g : g * wc;
                for (int u = 0; u < wc; u++)
                    idx[(size_t)(e * rec + off + u)] = h[2 * u] | h[2 * u + 1] << 16;
            }
    });
    mark("index chunks");
    const int64_t nfx = (int64_t)fx_rows.size(), npart = fx_off.back();
    // one entry of padding: the block kernel clamps its row-table loads to
    // the block's last entry, or to the first when a block touches no row
    rows.push_back(0);
    ppos.push_back(0);
    rs.push_back(0);
    for (int64_t i = 0; i < n; i++)
        if (bq[(size_t)i])
            fx_rows.push_back((int32_t)i);
    const int64_t ndir = (int64_t)fx_rows.size() - nfx;
    L.nelem = ne;
    L.n = n;
    L.nblk = nblk;
    L.nent = nent;
    L.nint = nint;
    L.nfx = nfx;
    L.ndir = ndir;
    L.npart = npart;
    if ((rc = put((void **)&L.eb_elem, eb_elem.data(), eb_elem.size() * 4)) ||
        (rc = put((void **)&L.eb_row, eb_row.data(), eb_row.size() * 4)) ||
        (rc = put((void **)&L.eb_nint, eb_nint.data(), eb_nint.size() * 4)) ||
        (rc = put((void **)&L.rows, rows.data(), rows.size() * 4)) ||
        (rc = put((void **)&L.rs, rs.data(), rs.size() * 2)) ||
        (rc = put((void **)&L.ppos, ppos.data(), ppos.size() * 4)) ||
        (rc = put((void **)&L.idx, idx.data(), idx.size() * 4)) ||
        (rc = put((void **)&L.fx_rows, fx_rows.data(), fx_rows.size() * 4)) ||
        (rc = put((void **)&L.fx_off, fx_off.data(), fx_off.size() * 4)) ||
        (rc = put((void **)&L.eperm, eperm.data(), (size_t)ne * 4)))
        return rc;
    mark("uploads");
    // per-row codes: partial start | count << 24 (when every start and count
    // fits)
    bool fits = codes && npart < ((int64_t)1 << 24) - 1;
    for (int64_t t = 0; t < nfx && fits; t++)
        fits = fx_off[(size_t)t + 1] - fx_off[(size_t)t] < 256;
    if (fits) {
        std::vector<uint32_t> code((size_t)n, EOP_NOT_SHARED);
        for (int64_t t = 0; t < nfx; t++)
            code[(size_t)fx_rows[(size_t)t]] =
                (uint32_t)fx_off[(size_t)t] |
                (uint32_t)(fx_off[(size_t)t + 1] - fx_off[(size_t)t]) << 24;
        if ((rc = put((void **)&L.fxc, code.data(), (size_t)n * 4)))
            return rc;
    }
    mark("shared-row codes");
    return finish();
    };

    std::shared_ptr<EopLayout> lay;
    if (same_space && sp->eop_layout && sp->eop_layout->n == n && sp->eop_layout->p == p &&
        sp->eop_layout_codes == codes) {
        lay = sp->eop_layout;
        mark("layout of the space");
    }
    else {
        try {
            lay = std::make_shared<EopLayout>();
            rc = build_layout(*lay);
        }
        catch (const std::bad_alloc &) {
            rc = set_error(PHG_HIP_E_NOMEM, "element operator set-up: out of host memory");
        }
        if (rc)
            return rc;      // lay frees what was allocated
        if (same_space) {
            sp->eop_layout = lay;
            sp->eop_layout_codes = codes;
        }
    }
    // per matrix: the coefficients and the partial sums
    EopData *eo = new EopData();
    eo->p = lay->p;
    eo->nb = lay->nb;
    eo->ks = lay->ks;
    eo->nelem = lay->nelem;
    eo->n = lay->n;
    eo->nblk = lay->nblk;
    eo->nent = lay->nent;
    eo->nint = lay->nint;
    eo->eb_elem = lay->eb_elem;
    eo->eb_row = lay->eb_row;
    eo->eb_nint = lay->eb_nint;
    eo->rows = lay->rows;
    eo->rs = lay->rs;
    eo->ppos = lay->ppos;
    eo->idx = lay->idx;
    eo->mass = pb->mass_coef != 0.0;
    eo->tiles = eo->mass ? lay->tiles_mass : lay->tiles;
    eo->msum = lay->msum;
    eo->npart = lay->npart;
    eo->fx_rows = lay->fx_rows;
    eo->fx_off = lay->fx_off;
    eo->nfx = lay->nfx;
    eo->ndir = lay->ndir;
    eo->fxc = lay->fxc;
    eo->grid_b = lay->grid_b;
    if (eo->mass) {
        const int gm = lay->p == 1 ? grid_block<1, true>()
                                   : (lay->p == 2 ? grid_block<2, true>() : grid_block<3, true>());
        eo->grid_b = (int)std::min<int64_t>(eo->grid_b, std::max(gm, 1));
    }
    eo->grid_f = lay->grid_f;
    eo->lay = lay;
    A->eop = eo;
    double *dk = nullptr;
    const int nc = eo->mass ? kNc<true> : kNc<false>;
    auto fail = [&](int code) {
        dfree(dk);
        free_eop(A);
        return code;
    };
    if ((rc = dmalloc((void **)&eo->part, (size_t)std::max<int64_t>(lay->npart, 1) * 4)) ||
        (rc = dmalloc((void **)&eo->coef, ((size_t)ne * nc + 2) * 4)))
        return fail(rc);
    if (pb->use_kappa &&
        ((rc = dmalloc((void **)&dk, (size_t)ne * 8)) ||
         (rc = upload(dk, kappa, (size_t)ne * 8))))
        return fail(rc);
    if ((rc = eop_coefficients(sp, pb->use_kappa ? dk : nullptr, lay->eperm, pb->mass_coef,
                               eo->coef, s)))
        return fail(rc);
    // two entries of padding, read (and multiplied by zero) for the last
    // element
    if (hipMemsetAsync(eo->coef + ne * nc, 0, 2 * sizeof(float), s) != hipSuccess)
        return fail(hip_fail(hipGetLastError(), "element operator set-up", __FILE__, __LINE__));
    hipError_t e = hipStreamSynchronize(s);
    dfree(dk);
    dk = nullptr;
    if (e != hipSuccess)
        return fail(hip_fail(e, "element operator set-up", __FILE__, __LINE__));
    mark("coefficients");
    eo->on = 1;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_mat_use_element_operator(phg_hip_mat *A, int32_t on)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null matrix");
    if (on && A->eop == nullptr)
        return phg::set_error(PHG_HIP_E_STATE,
                              "phg_hip_mat_use_element_operator: operator not built");
    if (A->eop)
        A->eop->on = on ? 1 : 0;
    return PHG_HIP_OK;
}
