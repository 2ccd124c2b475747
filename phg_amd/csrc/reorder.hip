// Solver numbering: the eliminated matrix renumbered so that DOFs follow a
// Hilbert curve of their nodes.
//
// PHG numbers DOFs vertex block first, then edges, then faces
// (dof.c:4821 phgDofMapE2D_), so a vertex row gathers x entries from the far
// end of the vector and every XCD's L2 sees most of x.  In the curve
// numbering a row's columns lie close to the row itself, and consecutive row
// blocks -- which the SpMV hands to the same XCD -- touch a compact window of
// x.  The renumbering is symmetric (rows and columns), keeps each row's
// entries in PHG's storage order, and is invisible to callers: the solvers
// gather b and x into the solver numbering on entry and scatter x back.
#include "phg_hip_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

namespace {

constexpr int TB = 256;

__global__ void k_morton(int64_t n, const double *xyz, double x0, double y0,
                         double z0, double sx, double sy, double sz,
                         uint64_t *key, int32_t *idx)
{
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n)
        return;
    const double m = 2097151.0;   // 2^21 - 1
    double fx = (xyz[3 * i] - x0) * sx, fy = (xyz[3 * i + 1] - y0) * sy,
           fz = (xyz[3 * i + 2] - z0) * sz;
    uint64_t qx = (uint64_t)fmin(fmax(fx * m, 0.0), m);
    uint64_t qy = (uint64_t)fmin(fmax(fy * m, 0.0), m);
    uint64_t qz = (uint64_t)fmin(fmax(fz * m, 0.0), m);
    key[i] = phg::hilbert3((uint32_t)qx, (uint32_t)qy, (uint32_t)qz);
    idx[i] = (int32_t)i;
}

// row permutation of a dim-valued space from the scalar node order
__global__ void k_expand_perm(int64_t ndof, int dim, const int32_t *pnode,
                              int32_t *perm, int32_t *iperm)
{
    int64_t k = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (k >= ndof)
        return;
    int64_t o = pnode[k];
    for (int c = 0; c < dim; c++) {
        perm[k * dim + c] = (int32_t)(o * dim + c);
        iperm[o * dim + c] = (int32_t)(k * dim + c);
    }
}

__global__ void k_perm_lengths(int64_t n, const int32_t *rp,
                               const int32_t *perm, int32_t *len)
{
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n)
        return;
    int32_t r = perm[i];
    len[i] = rp[r + 1] - rp[r];
}

// one group of PG lanes per row: the row's entries move in coalesced runs
constexpr int PG = 16, PRB = TB / PG;

__global__ void __launch_bounds__(TB) k_perm_fill(int64_t n, const int32_t *rp, const int32_t *ci,
                                                  const double *va, const double *diag,
                                                  const int32_t *perm, const int32_t *iperm,
                                                  const int64_t *qrp, int32_t *qci, double *qva,
                                                  double *qdiag)
{
    const int64_t i = (int64_t)blockIdx.x * PRB + threadIdx.x / PG;
    const int sl = threadIdx.x % PG;
    if (i >= n)
        return;
    const int32_t r = perm[i], b = rp[r], len = rp[r + 1] - b;
    const int64_t d = qrp[i];
    for (int32_t j = sl; j < len; j += PG) {
        qci[d + j] = iperm[ci[b + j]];
        qva[d + j] = va[b + j];
    }
    if (sl == 0)
        qdiag[i] = diag[r];
}

__global__ void k_narrow(int64_t n, const int64_t *a, int32_t *b)
{
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i <= n)
        b[i] = (int32_t)a[i];
}

inline unsigned nb(int64_t n) { return (unsigned)((n + TB - 1) / TB); }

}  // namespace

namespace phg {

void free_solver_copy(phg_hip_mat *A)
{
    free_dict(A);
    free_bsr(A);
    free_eop(A);       // built in this numbering
    if (A->q_owns) {
        dfree(A->q_rowptr);
        dfree(A->q_cols);
        dfree(A->q_vals);
        dfree(A->q_diag);
    }
    dfree(A->q_perm);
    dfree(A->q_iperm);
    dfree(A->q_vals32);
    dfree(A->q_diag32);
    dfree(A->q_rblk);
    dfree(A->q_rblk1k);
    dfree(A->q_wblk);
    A->q_perm = A->q_iperm = nullptr;
    A->q_rowptr = A->q_cols = nullptr;
    A->q_vals = A->q_diag = nullptr;
    A->q_vals32 = A->q_diag32 = nullptr;
    A->q_rblk = A->q_rblk1k = A->q_wblk = nullptr;
    A->q_nrblk = A->q_nrblk1k = A->q_nwblk = 0;
    A->q_owns = 0;
}

int ensure_solver_copy(phg_hip_mat *A)
{
    if (!A->q_stale)
        return PHG_HIP_OK;
    // the element operator is kept: its numbering is the node curve order
    // the rebuild derives again from the same coordinates, and its
    // coefficients were set with the values that made the PHG-order arrays
    EopData *eop = A->eop;
    A->eop = nullptr;
    int rc = build_solver_copy(A);
    A->eop = eop;
    if (rc == PHG_HIP_OK)
        A->q_stale = 0;
    // an operator reads the FP32 diagonal of its Dirichlet rows
    if (rc == PHG_HIP_OK && eop != nullptr)
        rc = ensure_fp32_copy(A);
    return rc;
}

int build_solver_copy(phg_hip_mat *A)
{
    hipStream_t s = ctx()->stream;
    const int64_t n = A->n;
    int rc;
    PhaseTrace tr("build_solver_copy");
    free_solver_copy(A);
    if (!A->reorder || A->d_xyz == nullptr || A->ndof * A->dim != n) {
        // solve in PHG's numbering
        A->q_rowptr = A->d_rowptr;
        A->q_cols = A->d_cols;
        A->q_vals = A->d_vals;
        A->q_diag = A->d_diag;
    }
    else {
        const int64_t nd = A->ndof;
        // bounding box of the nodes: the mesh's, kept from the space, or a
        // host pass over the coordinates
        double lo[3], hi[3];
        if (A->xyz_box) {
            for (int c = 0; c < 3; c++) {
                lo[c] = A->xyz_lo[c];
                hi[c] = A->xyz_hi[c];
            }
        }
        else {
            std::vector<double> xyz((size_t)nd * 3);
            if ((rc = download(xyz.data(), A->d_xyz, xyz.size() * sizeof(double))))
                return rc;
            for (int c = 0; c < 3; c++)
                lo[c] = hi[c] = xyz[(size_t)c];
            for (int64_t i = 0; i < nd; i++)
                for (int c = 0; c < 3; c++) {
                    lo[c] = std::min(lo[c], xyz[(size_t)(3 * i + c)]);
                    hi[c] = std::max(hi[c], xyz[(size_t)(3 * i + c)]);
                }
        }
        // one scale for all axes keeps the curve isotropic
        double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        double sc = ext > 0 ? 1.0 / ext : 1.0;
        uint64_t *key = nullptr, *key2 = nullptr;
        int32_t *idx = nullptr, *idx2 = nullptr, *iperm = nullptr, *len = nullptr;
        int64_t *qrp64 = nullptr;
        void *tmp = nullptr;
        size_t tmp_bytes = 0;
        auto cleanup = [&]() {
            dfree(key); dfree(key2); dfree(idx); dfree(idx2); dfree(iperm);
            dfree(len); dfree(qrp64); dfree(tmp);
        };
        if ((rc = dmalloc((void **)&key, (size_t)nd * 8)) ||
            (rc = dmalloc((void **)&key2, (size_t)nd * 8)) ||
            (rc = dmalloc((void **)&idx, (size_t)nd * 4)) ||
            (rc = dmalloc((void **)&idx2, (size_t)nd * 4))) {
            cleanup();
            return rc;
        }
        hipLaunchKernelGGL(k_morton, dim3(nb(nd)), dim3(TB), 0, s, nd,
                           (const double *)A->d_xyz, lo[0], lo[1], lo[2], sc,
                           sc, sc, key, idx);
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, key2,
                                               idx, idx2, (int)nd, 0, 63,
                                               s) != hipSuccess ||
            (rc = dmalloc(&tmp, tmp_bytes)) ||
            hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key2, idx,
                                               idx2, (int)nd, 0, 63,
                                               s) != hipSuccess) {
            cleanup();
            return rc ? rc : set_error(PHG_HIP_E_HIP, "radix sort failed");
        }
        tr.mark("curve keys, sort");
        if ((rc = dmalloc((void **)&A->q_perm, (size_t)n * 4)) ||
            (rc = dmalloc((void **)&iperm, (size_t)n * 4)) ||
            (rc = dmalloc((void **)&len, (size_t)(n + 1) * 4)) ||
            (rc = dmalloc((void **)&qrp64, (size_t)(n + 1) * 8))) {
            cleanup();
            return rc;
        }
        hipLaunchKernelGGL(k_expand_perm, dim3(nb(nd)), dim3(TB), 0, s, nd,
                           A->dim, (const int32_t *)idx2, A->q_perm, iperm);
        hipLaunchKernelGGL(k_perm_lengths, dim3(nb(n)), dim3(TB), 0, s, n,
                           (const int32_t *)A->d_rowptr,
                           (const int32_t *)A->q_perm, len);
        PHG_HIP_CHECK(hipMemsetAsync(len + n, 0, 4, s));
        size_t sbytes = 0;
        dfree(tmp);
        tmp = nullptr;
        if (hipcub::DeviceScan::ExclusiveSum(nullptr, sbytes, len, qrp64,
                                             (int)(n + 1), s) != hipSuccess ||
            (rc = dmalloc(&tmp, sbytes)) ||
            hipcub::DeviceScan::ExclusiveSum(tmp, sbytes, len, qrp64,
                                             (int)(n + 1), s) != hipSuccess) {
            cleanup();
            return rc ? rc : set_error(PHG_HIP_E_HIP, "scan failed");
        }
        A->q_owns = 1;
        if ((rc = dmalloc((void **)&A->q_rowptr, (size_t)(n + 1) * 4)) ||
            (rc = dmalloc((void **)&A->q_cols, (size_t)(A->nnz + 1) * 4)) ||
            (rc = dmalloc((void **)&A->q_vals, (size_t)(A->nnz + 1) * 8)) ||
            (rc = dmalloc((void **)&A->q_diag, (size_t)n * 8))) {
            cleanup();
            return rc;
        }
        hipLaunchKernelGGL(k_perm_fill, dim3((unsigned)((n + PRB - 1) / PRB)), dim3(TB), 0, s, n,
                           (const int32_t *)A->d_rowptr,
                           (const int32_t *)A->d_cols, (const double *)A->d_vals,
                           (const double *)A->d_diag, (const int32_t *)A->q_perm,
                           (const int32_t *)iperm, (const int64_t *)qrp64,
                           A->q_cols, A->q_vals, A->q_diag);
        hipLaunchKernelGGL(k_narrow, dim3(nb(n + 1)), dim3(TB), 0, s, n,
                           (const int64_t *)qrp64, A->q_rowptr);
        PHG_HIP_CHECK(hipGetLastError());
        PHG_HIP_CHECK(hipStreamSynchronize(s));
        tr.mark("renumbered copy");
        cleanup();
        tr.mark("frees");
    }
    // the 2048-entry row blocks of the solver copy (the FP64 residuals of
    // the mixed solvers and the B2K layouts use them); the 4096-entry and
    // wavefront blocks are split when a layout first needs them
    if ((rc = split_row_blocks(A->q_rowptr, n, A->nnz, SPMV_BLOCK_NNZ, SPMV_BLOCK_ROWS,
                               &A->q_rblk, &A->q_nrblk)))
        return rc;
    tr.mark("row-block split");
    return PHG_HIP_OK;
}

}  // namespace phg
