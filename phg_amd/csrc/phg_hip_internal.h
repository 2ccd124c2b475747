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
    // FP32 copy of the values for the mixed-precision inner solve
    float *d_vals32 = nullptr;
    float *d_diag32 = nullptr;
    // SpMV row blocks: consecutive rows whose entries fit the LDS stage of
    // the streaming SpMV (or one long row); boundaries [nrblk + 1]
    int32_t *d_rblk = nullptr;
    int64_t nrblk = 0;
};

namespace phg {
// Krylov building blocks (krylov.hip) used by the C ABI layer.
int spmv64(const phg_hip_mat *A, const double *x, double *y, double alpha,
           double beta, hipStream_t s);
int build_row_blocks(phg_hip_mat *A);
int ensure_fp32_copy(phg_hip_mat *A);
int setup_diagonal(phg_hip_mat *A);
}  // namespace phg
