// Internal declarations shared by the HIP translation units of libphg_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/phg_hip.h"

namespace phg {

// error state ------------------------------------------------------------
int set_error(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define PHG_HIP_CHECK(call)                                                   \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::phg::hip_fail(e_, #call, __FILE__, __LINE__);            \
    } while (0)

// library context ----------------------------------------------------------
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    // pinned host mailbox for solver state read-back
    void *host_mailbox = nullptr;
};
Context *ctx();                // nullptr if not initialised
int require_ctx();             // PHG_HIP_OK or PHG_HIP_E_STATE

// device memory helpers (stream-ordered frees are not used: every buffer is
// owned by a handle and freed on destroy)
int dmalloc(void **p, size_t bytes);
void dfree(void *p);
int upload(void *d, const void *h, size_t bytes);
int download(void *h, const void *d, size_t bytes);

}  // namespace phg

// opaque handles -----------------------------------------------------------
struct phg_hip_space {
    int32_t order, nbas;
    int64_t nvert, nelem, ndof;
    double *d_verts = nullptr;       // [nvert][3]
    int32_t *d_elem_verts = nullptr; // [nelem][4]
    int32_t *d_map = nullptr;        // [nelem][nbas]
    uint8_t *d_bdry = nullptr;       // [ndof]
};

struct phg_hip_vec {
    int64_t n;
    double *d = nullptr;
};

struct phg_hip_mat {
    int64_t n = 0, nnz = 0;
    int32_t *d_rowptr = nullptr;     // [n+1] (nnz < 2^31 is checked)
    int32_t *d_cols = nullptr;       // [nnz]
    double *d_vals = nullptr;        // [nnz]
    double *d_rhs = nullptr;         // [n]
    uint8_t *d_bdry_row = nullptr;   // [n]
    double *d_diag = nullptr;        // [n], set after elimination
    int eliminated = 0;
    // SpMV row blocks: consecutive rows whose entries fit the LDS stage of
    // the streaming SpMV (or one long row); boundaries [nrblk + 1]
    int32_t *d_rblk = nullptr;
    int64_t nrblk = 0;
    // the same for the wavefront-level stream (smaller blocks)
    int32_t *d_wblk = nullptr;
    int64_t nwblk = 0;
    // FP32 copies of the CSR values and diagonal (mixed precision)
    float *d_vals32 = nullptr;
    float *d_diag32 = nullptr;
    // SpMV storage format used by the solvers, per precision, and the times
    // the autotuner measured for each format [fmt][0: FP64, 1: FP32] (us)
    int fmt64 = 1, fmt32 = 1;
    double tune_us[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    // SELL-64 copy used by the solvers: rows sorted by length inside windows
    // of SELL_WINDOW rows and renumbered symmetrically (perm: new -> old),
    // slices of 64 rows stored column-major, padded with zero entries
    int64_t nslice = 0;
    int32_t *d_perm = nullptr, *d_iperm = nullptr;
    int64_t *d_sptr = nullptr;       // [nslice + 1] first entry of each slice
    int32_t *d_swidth = nullptr;     // [nslice] entries per row in the slice
    int32_t *d_scols = nullptr;
    double *d_svals = nullptr;
    float *d_svals32 = nullptr;
    double *d_sdiag = nullptr;       // permuted diagonal
    float *d_sdiag32 = nullptr;
    int64_t sell_nnz = 0;            // stored entries including padding
};

namespace phg {
// Krylov building blocks (krylov.hip) used by the C ABI layer.
int spmv64(const phg_hip_mat *A, const double *x, double *y, double alpha,
           double beta, hipStream_t s);
int build_row_blocks(phg_hip_mat *A);
int build_sell(phg_hip_mat *A);
void free_sell(phg_hip_mat *A);
int ensure_fp32_copy(phg_hip_mat *A);
int setup_diagonal(phg_hip_mat *A);
}  // namespace phg
