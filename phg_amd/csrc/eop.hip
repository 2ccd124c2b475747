// Matrix-free FP32 operator of a scalar stiffness matrix, element products on
// the matrix cores.
//
// For an affine tetrahedron the stiffness of quad.c:5260
// (phgQuadGradBasAGradBas with a P0 coefficient) factors as
//     K_e = sum_{m=0..5} c_m(e) T^m,
//     c_m(e) = kappa_e vol_e (grad lambda_a . grad lambda_b),  m = (a <= b),
// over the independent barycentric coordinates a, b = 1..3, with reference
// tensors T^m fixed per order (assemble.hip computes both).  Applying the
// operator to x is then, per element,
//     y_e[i] = sum_m c_m(e) (T^m x_e)[i],
// and T^m x_e for 16 elements at a time is a matrix product: rows (i, m) of
// the stacked T against the 16 gathered element vectors.  It runs as
// v_mfma_f32_16x16x32_bf16 with FP32 precision kept by splitting both
// operands into three bf16 parts (t = t0 + t1 + t2, x = x0 + x1 + x2, exact
// for FP32 x) and accumulating the six products whose order is at most
// 2^-16 relative (t0x0, t0x1, t1x0, t0x2, t1x1, t2x0) in FP32; the dropped
// terms are below 2^-24.  One 16-row output tile holds the slots m (8 rows,
// 6 used) of two local indices i = 2t, 2t + 1 for 16 elements; the A
// fragments (T parts) of all tiles stay in LDS for the kernel, the B
// fragments (x parts) are written to LDS per 16 elements.
//
// The per-row sums y_i = sum over the row's incident (element, local index)
// pairs, in element order, run in krylov.hip (k_eop_gather) with the
// solver's epilogue.  Against the assembled FP32 CSR the operator moves
// map + coefficients + element results (P3: 80 + 24 + 2 x 80 B per element
// plus 4 B per incidence) instead of about 1.4 KB of matrix entries per
// element.
#include "phg_hip_internal.h"

#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int EB = 512;           // work-items per work-group (8 wavefronts)
constexpr int EW = EB / 64;
constexpr int NC = 6;             // coefficients per element

template <int P> struct EopShape {
    static constexpr int NB = P == 1 ? 4 : (P == 2 ? 10 : 20);
    static constexpr int NT = NB / 2;                 // output tiles
    static constexpr int KS = (6 * NB + 31) / 32;     // K steps of 32
    static constexpr int KP = KS * 32 + 8;            // B row stride (bank skew)
};

// A fragment of (tile t, K step s) for lane l: row R = l % 16 is slot
// m = R % 8 of local index i = 2 t + R / 8; K indices 32 s + 8 (l / 16) + u,
// u = 0..7; K index k = q NB + j is product q of the split, column j of
// T^m (row i).
constexpr int kTPart[6] = {0, 0, 1, 0, 1, 2};
constexpr int kXPart[6] = {0, 1, 0, 2, 1, 0};

// element tiles of 16 a wavefront takes per round: their gathers are
// issued together and every A fragment read from LDS feeds ET products
template <int P> constexpr int kEt = P == 3 ? 2 : 4;

template <int P>
__global__ void __launch_bounds__(EB)
k_eop_elem(int64_t nelem, const int32_t *__restrict__ map,
           const float *__restrict__ coef, const uint4 *__restrict__ tiles,
           const float *__restrict__ x, float *__restrict__ y,
           const int32_t *done, const int32_t *mode, int32_t want)
{
    using S = EopShape<P>;
    constexpr int NB = S::NB, NT = S::NT, KS = S::KS, KP = S::KP, ET = kEt<P>;
    constexpr int NJ = (NB + 3) / 4;      // local indices j per lane
    if (done != nullptr && (*done != 0 || *mode != want))
        return;
    __shared__ uint4 at[NT * KS * 64];
    __shared__ __align__(16) __bf16 bt[EW][16][KP];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int n = l & 15, g = l >> 4;
    for (int t = tid; t < NT * KS * 64; t += EB)
        at[t] = tiles[t];
    // K padding of the B rows stays zero
    constexpr int PADW = KP - 6 * NB;
    for (int t = l; t < 16 * PADW; t += 64)
        bt[w][t / PADW][6 * NB + t % PADW] = (__bf16)0.0f;
    const int64_t ntile = (nelem + 15) / 16;
    // every wavefront of the work-group runs the same number of rounds, so
    // the work-group barriers below are reached by all of them
    for (int64_t base = (int64_t)blockIdx.x * EW * ET; base < ntile;
         base += (int64_t)gridDim.x * EW * ET) {
        const int64_t tile0 = base + (int64_t)w * ET;
        // gathers of the ET tiles: x of local indices j = g, g + 4, ... and
        // the coefficients of the D rows 4 g + r (slots m = 4 (g % 2) + r)
        float xv[ET][NJ], c[ET][4];
#pragma unroll
        for (int et = 0; et < ET; et++) {
            const int64_t e = (tile0 + et) * 16 + n;
            const bool ok = tile0 + et < ntile && e < nelem;
#pragma unroll
            for (int jj = 0; jj < NJ; jj++) {
                const int j = 4 * jj + g;
                int32_t col = -1;
                if (ok && j < NB)
                    col = map[e * NB + j];
                xv[et][jj] = col >= 0 ? x[col] : 0.0f;
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = 4 * (g & 1) + r;
                c[et][r] = (ok && m < NC) ? coef[e * NC + m] : 0.0f;
            }
        }
        // split into bf16 parts through the wavefront's LDS rows, one tile
        // at a time, into B fragments
        bf16x8 b[ET][KS];
#pragma unroll
        for (int et = 0; et < ET; et++) {
            __syncthreads();    // the previous fragments are read
#pragma unroll
            for (int jj = 0; jj < NJ; jj++) {
                const int j = 4 * jj + g;
                if (j < NB) {
                    const float v = xv[et][jj];
                    const __bf16 x0 = (__bf16)v;
                    const float r1 = v - (float)x0;
                    const __bf16 x1 = (__bf16)r1;
                    const __bf16 x2 = (__bf16)(r1 - (float)x1);
                    const __bf16 xp[3] = {x0, x1, x2};
#pragma unroll
                    for (int q = 0; q < 6; q++)
                        bt[w][n][q * NB + j] = xp[kXPart[q]];
                }
            }
            __syncthreads();
#pragma unroll
            for (int s2 = 0; s2 < KS; s2++)
                b[et][s2] = *(const bf16x8 *)&bt[w][n][32 * s2 + 8 * g];
        }
        float yv[ET][(NT + 1) / 2];
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
                    acc[et] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[et][s2], acc[et],
                                                                      0, 0, 0);
            }
            // contract with c_m, then add the two row groups of local index
            // 2 t + g / 2 (lanes g and g ^ 1)
#pragma unroll
            for (int et = 0; et < ET; et++) {
                float v = c[et][0] * acc[et][0] + c[et][1] * acc[et][1] +
                          c[et][2] * acc[et][2] + c[et][3] * acc[et][3];
                v += __shfl_xor(v, 16, 64);
                if ((t & 1) == (g & 1))
                    yv[et][t >> 1] = v;
            }
        }
#pragma unroll
        for (int et = 0; et < ET; et++) {
            const int64_t e = (tile0 + et) * 16 + n;
            if (tile0 + et < ntile && e < nelem) {
#pragma unroll
                for (int t = 0; t < NT; t++)
                    if ((t & 1) == (g & 1))
                        y[e * NB + 2 * t + (g >> 1)] = yv[et][t >> 1];
            }
        }
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

// A fragments [nb / 2][ks][64][8] of the stacked split T (see above)
void build_tiles(int p, int nb, int ks, std::vector<uint16_t> &out)
{
    std::vector<double> T((size_t)NC * nb * nb);
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
                    if (m < NC && q < 6)
                        v = part[(size_t)kTPart[q] * T.size() +
                                 ((size_t)m * nb + i) * nb + j];
                    out[(((size_t)t * ks + s) * 64 + l) * 8 + u] = v;
                }
}

uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

int grid_elem(int p)
{
    const void *k = p == 1 ? (const void *)k_eop_elem<1>
                  : (p == 2 ? (const void *)k_eop_elem<2> : (const void *)k_eop_elem<3>);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, EB, 0) != hipSuccess ||
        per_cu <= 0) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const int cus = phg::ctx()->num_cus > 0 ? phg::ctx()->num_cus : 256;
    return per_cu * cus;
}

}  // namespace

namespace phg {

void free_eop(phg_hip_mat *A)
{
    EopData *eo = A->eop;
    if (eo == nullptr)
        return;
    dfree(eo->map);
    dfree(eo->coef);
    dfree(eo->tiles);
    dfree(eo->inc_off);
    dfree(eo->inc);
    dfree(eo->gblk);
    dfree(eo->bdry);
    dfree(eo->y);
    delete eo;
    A->eop = nullptr;
}

void eop_elements(const EopData *eo, const float *x, const int32_t *done,
                  const int32_t *mode, int32_t want, hipStream_t s)
{
    if (eo->nelem == 0)
        return;
    const int grid = eo->grid_e;
    const uint4 *t = (const uint4 *)eo->tiles;
    if (eo->p == 1)
        hipLaunchKernelGGL(k_eop_elem<1>, dim3(grid), dim3(EB), 0, s, eo->nelem,
                           (const int32_t *)eo->map, (const float *)eo->coef, t, x,
                           eo->y, done, mode, want);
    else if (eo->p == 2)
        hipLaunchKernelGGL(k_eop_elem<2>, dim3(grid), dim3(EB), 0, s, eo->nelem,
                           (const int32_t *)eo->map, (const float *)eo->coef, t, x,
                           eo->y, done, mode, want);
    else
        hipLaunchKernelGGL(k_eop_elem<3>, dim3(grid), dim3(EB), 0, s, eo->nelem,
                           (const int32_t *)eo->map, (const float *)eo->coef, t, x,
                           eo->y, done, mode, want);
}

}  // namespace phg

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
    if (!A->eliminated || A->dim != 1 || pb->dim != 1 || pb->mass_coef != 0.0 ||
        A->n != sp->ndof || sp->order < 1 || sp->order > 3 ||
        (pb->use_kappa && kappa == nullptr))
        return set_error(PHG_HIP_E_ARG,
                         "phg_hip_mat_element_operator: needs the eliminated matrix "
                         "of a scalar stiffness problem (no mass term) on this space");
    if ((rc = ensure_fp32_copy(A)))
        return rc;
    free_eop(A);
    hipStream_t s = ctx()->stream;
    const int p = sp->order, nb = sp->nbas;
    const int64_t ne = sp->nelem, n = A->n;
    if ((int64_t)nb * ne > INT32_MAX)
        return set_error(PHG_HIP_E_ARG, "phg_hip_mat_element_operator: too many elements");
    std::vector<int32_t> map((size_t)ne * nb), ev((size_t)ne * 4), perm;
    std::vector<double> xyz((size_t)sp->nvert * 3);
    std::vector<uint8_t> brow((size_t)n);
    if ((rc = download(map.data(), sp->d_map, map.size() * 4)) ||
        (rc = download(ev.data(), sp->d_elem_verts, ev.size() * 4)) ||
        (rc = download(xyz.data(), sp->d_verts, xyz.size() * 8)) ||
        (rc = download(brow.data(), A->d_bdry_row, (size_t)n)))
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
    // elements in Morton order of their centroids
    std::vector<int32_t> eperm((size_t)ne);
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t v = 0; v < sp->nvert; v++)
            for (int c = 0; c < 3; c++) {
                lo[c] = std::min(lo[c], xyz[(size_t)(3 * v + c)]);
                hi[c] = std::max(hi[c], xyz[(size_t)(3 * v + c)]);
            }
        double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        const double sc = ext > 0 ? 2097151.0 / ext : 0.0;
        std::vector<uint64_t> key((size_t)ne);
        for (int64_t e = 0; e < ne; e++) {
            uint64_t q[3];
            for (int c = 0; c < 3; c++) {
                double m = 0.0;
                for (int v = 0; v < 4; v++)
                    m += xyz[(size_t)(3 * ev[(size_t)(4 * e + v)] + c)];
                double f = (0.25 * m - lo[c]) * sc;
                q[c] = (uint64_t)std::min(std::max(f, 0.0), 2097151.0);
            }
            key[(size_t)e] = spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2;
        }
        std::iota(eperm.begin(), eperm.end(), 0);
        std::stable_sort(eperm.begin(), eperm.end(),
                         [&](int32_t a, int32_t b) { return key[(size_t)a] < key[(size_t)b]; });
    }
    // element DOF map in the solver numbering, Dirichlet DOFs -1; row
    // incidence lists (non-Dirichlet rows) in element order
    std::vector<int32_t> emap((size_t)ne * nb), off((size_t)n + 1, 0), inc;
    std::vector<uint8_t> bq((size_t)n);
    for (int64_t i = 0; i < n; i++)
        bq[(size_t)i] = brow[(size_t)(A->q_perm ? perm[(size_t)i] : i)];
    for (int64_t e2 = 0; e2 < ne; e2++) {
        const int64_t e = eperm[(size_t)e2];
        for (int a = 0; a < nb; a++) {
            const int32_t r = iperm[(size_t)map[(size_t)(e * nb + a)]];
            const bool b = bq[(size_t)r] != 0;
            emap[(size_t)(e2 * nb + a)] = b ? -1 : r;
            if (!b)
                off[(size_t)r + 1]++;
        }
    }
    for (int64_t i = 0; i < n; i++)
        off[(size_t)i + 1] += off[(size_t)i];
    const int64_t ninc = off[(size_t)n];
    inc.resize((size_t)std::max<int64_t>(ninc, 1));
    {
        std::vector<int32_t> pos(off.begin(), off.end() - 1);
        for (int64_t e2 = 0; e2 < ne; e2++)
            for (int a = 0; a < nb; a++) {
                const int32_t r = emap[(size_t)(e2 * nb + a)];
                if (r >= 0)
                    inc[(size_t)pos[(size_t)r]++] = (int32_t)(e2 * nb + a);
            }
    }
    int ks = p == 1 ? EopShape<1>::KS : (p == 2 ? EopShape<2>::KS : EopShape<3>::KS);
    std::vector<uint16_t> tiles;
    build_tiles(p, nb, ks, tiles);
    EopData *eo = new EopData();
    eo->p = p;
    eo->nb = nb;
    eo->ks = ks;
    eo->nelem = ne;
    eo->n = n;
    eo->ninc = ninc;
    A->eop = eo;
    double *dk = nullptr;
    int32_t *dperm = nullptr;
    auto fail = [&](int code) {
        dfree(dk);
        dfree(dperm);
        free_eop(A);
        return code;
    };
    if ((rc = dmalloc((void **)&eo->map, emap.size() * 4)) ||
        (rc = dmalloc((void **)&eo->coef, (size_t)ne * NC * 4)) ||
        (rc = dmalloc((void **)&eo->tiles, tiles.size() * 2)) ||
        (rc = dmalloc((void **)&eo->inc_off, off.size() * 4)) ||
        (rc = dmalloc((void **)&eo->inc, inc.size() * 4)) ||
        (rc = dmalloc((void **)&eo->bdry, (size_t)n)) ||
        (rc = dmalloc((void **)&eo->y, (size_t)ne * nb * 4)) ||
        (rc = dmalloc((void **)&dperm, (size_t)ne * 4)) ||
        (rc = upload(eo->map, emap.data(), emap.size() * 4)) ||
        (rc = upload(eo->tiles, tiles.data(), tiles.size() * 2)) ||
        (rc = upload(eo->inc_off, off.data(), off.size() * 4)) ||
        (rc = upload(eo->inc, inc.data(), inc.size() * 4)) ||
        (rc = upload(eo->bdry, bq.data(), (size_t)n)) ||
        (rc = upload(dperm, eperm.data(), (size_t)ne * 4)))
        return fail(rc);
    if (pb->use_kappa &&
        ((rc = dmalloc((void **)&dk, (size_t)ne * 8)) ||
         (rc = upload(dk, kappa, (size_t)ne * 8))))
        return fail(rc);
    if ((rc = eop_coefficients(sp, pb->use_kappa ? dk : nullptr, dperm, eo->coef, s)) ||
        (rc = split_row_blocks(eo->inc_off, n, ninc, EOP_GATHER_NNZ, EOP_GATHER_ROWS,
                               &eo->gblk, &eo->ngblk)))
        return fail(rc);
    hipError_t e = hipStreamSynchronize(s);
    dfree(dk);
    dfree(dperm);
    dk = nullptr;
    dperm = nullptr;
    if (e != hipSuccess)
        return fail(hip_fail(e, "element operator set-up", __FILE__, __LINE__));
    const int et = p == 1 ? kEt<1> : (p == 2 ? kEt<2> : kEt<3>);
    const int64_t need = ((ne + 15) / 16 + (int64_t)EW * et - 1) / ((int64_t)EW * et);
    eo->grid_e = (int)std::max<int64_t>(1, std::min<int64_t>(need, grid_elem(p)));
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
