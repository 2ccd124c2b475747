// Internal declarations shared by the HIP translation units of libphg_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <memory>
#include <vector>

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

// SpMV storage formats (PHG_HIP_FMT_*) the per-matrix timing table holds
constexpr int MAX_FORMATS = 16;
static_assert(PHG_HIP_FMT_COUNT <= MAX_FORMATS, "timing table too small");

// library context ----------------------------------------------------------
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    // pinned host mailbox for solver state read-back
    void *host_mailbox = nullptr;
    // extra streams (default priority) for concurrent solves, made on demand
    static constexpr int NAUX = 8;
    hipStream_t aux[NAUX] = {};
    int naux = 0;
};
Context *ctx();                // nullptr if not initialised
int require_ctx();             // PHG_HIP_OK or PHG_HIP_E_STATE
int aux_streams(hipStream_t **pool, int *n);   // the extra stream pool

// device memory helpers (stream-ordered frees are not used: every buffer is
// owned by a handle and freed on destroy)
int dmalloc(void **p, size_t bytes);
void dfree(void *p);
int upload(void *d, const void *h, size_t bytes);
int download(void *h, const void *d, size_t bytes);
// per-device address of the scratch array of the no-allocation entry point
// (assemble.hip), taken by phg_hip_init with the device current
int noalloc_init(int device);

// Phase timer of the set-up paths (assembly, elimination, solver copy):
// with PHG_HIP_SETUP_TRACE set in the environment, mark(what) waits for the
// library stream and prints the time since the previous mark to stderr;
// otherwise it does nothing.
struct PhaseTrace {
    explicit PhaseTrace(const char *scope);
    void mark(const char *what);
    const char *scope;
    bool on;
    double t_last;
};

}  // namespace phg

// Element-block layout of the matrix-free operator (eop.hip, see EopData
// below): the element order, the blocks, their row tables, index chunks
// and the shared rows.  It depends on the space (element maps, vertex
// coordinates), its Dirichlet rows and the solver numbering, not on the
// coefficients, so the operators of all matrices assembled on one space
// share it.  The destructor frees the device arrays.
struct EopLayout {
    int p = 0, nb = 0, ks = 0;
    int64_t nelem = 0, n = 0, nblk = 0, nent = 0, nint = 0, nfx = 0, ndir = 0, npart = 0;
    int grid_b = 0, grid_f = 0;
    int32_t *eb_elem = nullptr, *eb_row = nullptr, *eb_nint = nullptr, *rows = nullptr;
    uint16_t *rs = nullptr;
    int32_t *ppos = nullptr;
    uint32_t *idx = nullptr;
    uint8_t *tiles = nullptr;        // A fragments of the stiffness tensors
    uint8_t *tiles_mass = nullptr;   // the same with the mass tensor in slot 6
    // mass tensor row sums: P1 / P3 per output tile (local index 2 t = 2 t + 1),
    // P2 per output row of the point-product path
    float msum[16] = {};
    int32_t *fx_rows = nullptr, *fx_off = nullptr;
    uint32_t *fxc = nullptr;
    int32_t *eperm = nullptr;        // [nelem] curve position -> element
    int on_device = 0;               // built on the device (fixed element blocks)
    ~EopLayout();
};

// Assembly workspace of a space (assemble.hip): the row incidence of the
// last problem dimension assembled (a function of the element maps alone)
// and the scratch buffers of phg_hip_assemble, kept between assemblies and
// freed with the space.
struct AsmWorkspace;
void free_asm_workspace(AsmWorkspace *w);

// opaque handles -----------------------------------------------------------
struct phg_hip_space {
    AsmWorkspace *asm_ws = nullptr;
    int32_t order, nbas;
    int64_t nvert, nelem, ndof;
    uint64_t uid = 0;                // unique per space of the process
    // element-operator layout of matrices assembled on this space (built by
    // the first phg_hip_mat_element_operator on one; codes: with the per-row
    // codes of the shared rows)
    mutable std::shared_ptr<EopLayout> eop_layout;
    mutable int eop_layout_codes = -1;
    double *d_verts = nullptr;       // [nvert][3]
    int32_t *d_elem_verts = nullptr; // [nelem][4]
    int32_t *d_map = nullptr;        // [nelem][nbas]
    uint8_t *d_bdry = nullptr;       // [ndof]
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};   // bounding box of the vertices
    int batch_refs = 0;              // batches whose graphs read its arrays
};

struct phg_hip_vec {
    int64_t n;
    double *d = nullptr;
    int batch_refs = 0;              // batches whose graphs read or write it
};

// Blocks of rows; per block a dictionary of the distinct column groups
// (column >> shift) its entries touch; every entry stores its column as 16
// bits: slot << shift | (column & (2^shift - 1)).  Entries of a block start
// at a multiple of 8 (padding: index 0, value 0) and each group of 512
// entries is stored transposed (8 per work-item).
struct DictLayout {
    int state = 0;                   // 0 not built, 1 built, -1 not representable
    int shift = 0, max_slots = 0;
    int64_t nblk = 0, nent = 0, ndict = 0;
    int32_t *rowptr = nullptr;       // [n+1] entry positions (padded)
    int32_t *blk_row = nullptr;      // [nblk+1] first row of each block
    int32_t *blk_ent = nullptr;      // [nblk+1] first entry of each block
    int32_t *blk_dict = nullptr;     // [nblk+1] first dictionary slot
    int32_t *dict = nullptr;         // group bases (multiples of 2^shift)
    uint16_t *idx = nullptr;
    float *val32 = nullptr;
    double *val64 = nullptr;
};

// Matrix-free FP32 operator of a scalar stiffness matrix (eop.hip), in the
// solver numbering: y = sum_e P_e^T K_e P_e x with
// K_e = sum_m coef[e][m] T^m (6 pairs m of the independent barycentric
// coordinates, T^m the reference tensors), the element products on bf16
// MFMA with FP32 operands
// split into three bf16 parts, then an ordered per-row sum over the row's
// incident (element, local index) pairs.  Elements are in Hilbert-curve order of
// their centroids; Dirichlet DOFs have map -1 (their columns are the ones
// elimination removes) and their rows are diag * x.
struct EopData {
    int p = 0, nb = 0, ks = 0;        // order, basis size, K steps of 32
    int on = 0;                       // used by the FP32 solver slots
    int64_t nelem = 0, n = 0;
    // element blocks (consecutive elements in curve order)
    int64_t nblk = 0, nent = 0, nint = 0;   // blocks, touched-row entries, interior rows
    int32_t *eb_elem = nullptr;       // [nblk+1] element ranges
    int32_t *eb_row = nullptr;        // [nblk+1] touched-row entry ranges
    int32_t *eb_nint = nullptr;       // [nblk] interior rows of each block
    int32_t *rows = nullptr;          // [nent] solver rows, interior first per block
    uint16_t *rs = nullptr;           // [nent+nblk] slot starts per block
    int32_t *ppos = nullptr;          // [nent] partial slot of a shared row
    uint32_t *idx = nullptr;          // [nelem][4][w] lane-group chunks of 16-bit row
                                      // indices and slots (0xffff: none)
    int mass = 0;                     // mass term: 7 coefficients per element
    float *coef = nullptr;            // [nelem][6 or 7] + 2 entries of padding
    uint8_t *tiles = nullptr;         // A fragments (bf16 parts for P3, FP32 for P1 / P2)
    const float *msum = nullptr;      // [16] mass row sums (layout's)
    float *part = nullptr;            // [npart] partial sums of shared rows
    int64_t npart = 0;
    // fix-up: nfx shared rows (partial ranges fx_off) then ndir Dirichlet rows
    int32_t *fx_rows = nullptr;       // [nfx + ndir]
    int32_t *fx_off = nullptr;        // [nfx + 1]
    int64_t nfx = 0, ndir = 0;
    uint32_t *fxc = nullptr;          // [n] partial start | count << 24 of a shared
                                      // row, EOP_NOT_SHARED otherwise (null: no codes)
    int grid_b = 0, grid_f = 0;       // launch grids; grid_b partial sums of (x, y)
    // the layout the pointers above (all but coef and part) belong to
    std::shared_ptr<EopLayout> lay;
};

struct phg_hip_mat {
    int64_t n = 0, nnz = 0;
    int dim = 1;
    // PHG-order CSR (matvec.c packed storage): phg_hip_mat_download and
    // phg_hip_matvec work on it
    int32_t *d_rowptr = nullptr;     // [n+1] (nnz < 2^31 is checked)
    int32_t *d_cols = nullptr;       // [nnz]
    double *d_vals = nullptr;        // [nnz]
    double *d_rhs = nullptr;         // [n]
    uint8_t *d_bdry_row = nullptr;   // [n]
    double *d_diag = nullptr;        // [n], set after elimination
    int eliminated = 0;
    // row blocks of the PHG-order CSR for the streaming SpMV
    int32_t *d_rblk = nullptr;
    int64_t nrblk = 0;
    // nodal coordinates of the scalar DOFs [ndof][3], kept from assembly to
    // derive the solver numbering
    double *d_xyz = nullptr;
    double xyz_lo[3] = {0, 0, 0}, xyz_hi[3] = {0, 0, 0};   // bounding box of the nodes
    int xyz_box = 0;                                       // xyz_lo/hi set
    int64_t ndof = 0;
    uint64_t space_uid = 0;          // uid of the space phg_hip_assemble used (0: none)

    // Solver copy: the eliminated matrix under a symmetric renumbering that
    // orders DOFs along a Hilbert curve of their nodes (q_perm: solver index ->
    // PHG index); the q_* arrays alias the PHG-order ones when q_perm is null
    int reorder = 1;                 // 1: derive the spatial numbering
    // 1: the PHG-order arrays changed (a batch re-assembled the matrix) and
    // the solver copy is rebuilt when next needed (ensure_solver_copy)
    int q_stale = 0;
    int32_t *q_perm = nullptr;
    int32_t *q_iperm = nullptr;      // PHG index -> solver index (batch views only)
    int q_owns = 0;
    int32_t *q_rowptr = nullptr, *q_cols = nullptr;
    double *q_vals = nullptr, *q_diag = nullptr;
    float *q_vals32 = nullptr, *q_diag32 = nullptr;
    // row blocks of SPMV_BLOCK_NNZ / SPMV_LARGE_NNZ / SPMV_WAVE_NNZ entries
    int32_t *q_rblk = nullptr, *q_rblk1k = nullptr, *q_wblk = nullptr;
    int64_t q_nrblk = 0, q_nrblk1k = 0, q_nwblk = 0;

    // SpMV storage format used by the solvers, per precision, and the times
    // the autotuner measured for each format [fmt][0: FP64, 1: FP32] (us)
    int fmt64 = 0, fmt32 = 0;
    double tune_us[phg::MAX_FORMATS][2] = {};
    // SELL-64 copy built from the solver CSR: rows sorted by length inside
    // windows of 4096 rows (d_perm: SELL index -> PHG index, composed with
    // q_perm), slices of 64 rows stored column-major, padded with zeros
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
    // Dictionary-compressed CSR copies of the solver matrix (dictcsr.hip):
    // cd with 256-column chunks (columns decoded through the dictionary),
    // xs with 16-column lines (the lines are staged in LDS by the SpMV)
    DictLayout cd, xs;
    // block copy for the many-right-hand-side solver (block.hip): 3x3
    // blocks for 3-component problems (b_dim 3), else the solver CSR itself
    // (b_dim 1, arrays aliased); b_dim 0: not built
    int b_dim = 0, b_owns = 0;
    int64_t b_nb = 0, b_nblk = 0;
    int32_t *b_ptr = nullptr, *b_col = nullptr;
    double *b_val64 = nullptr;
    float *b_val32 = nullptr;
    float *b_rsum32 = nullptr;       // [b_nb][b_dim^2] row-block sums (FP64 sums in FP32)
    // solver workspace of this matrix (krylov.hip)
    void *ws = nullptr;
    // SpMM wavefront shape index per precision (block.hip) and tuner times
    int spmm_shape[2] = {0, 0};
    double spmm_us[8][2] = {};
    // matrix-free FP32 operator (eop.hip), null unless built
    EopData *eop = nullptr;
    // batches (phg_hip_batch) whose graphs hold this matrix's buffers: while
    // nonzero, calls that would free them (destroy, a new element operator)
    // are refused
    int batch_refs = 0;
};

namespace phg {
// row-block sizes of the streaming SpMV kernels (entries / rows per block)
constexpr int SPMV_BLOCK_NNZ = 2048, SPMV_BLOCK_ROWS = 256;
constexpr int SPMV_WAVE_NNZ = 512, SPMV_WAVE_ROWS = 64;
constexpr int SPMV_LARGE_NNZ = 4096;
constexpr int DICT_BLOCK_NNZ = 4096, DICT_BLOCK_ROWS = 256;
constexpr int DICT_GROUP = 512;   // entries of one wavefront (64 x 8)
// Krylov building blocks (krylov.hip) used by the C ABI layer.
int spmv64(const phg_hip_mat *A, const double *x, double *y, double alpha,
           double beta, hipStream_t s);
void free_workspace(phg_hip_mat *A);
int build_row_blocks(phg_hip_mat *A);
// solver copy (reorder.hip)
int build_solver_copy(phg_hip_mat *A);
// rebuilds the solver copy when q_stale is set (every entry point that reads
// the q_* arrays or the element operator calls it first)
int ensure_solver_copy(phg_hip_mat *A);
void free_solver_copy(phg_hip_mat *A);
int split_row_blocks(const int32_t *d_rowptr, int64_t n, int64_t nnz,
                     int64_t cap, int64_t rcap, int32_t **d_blk, int64_t *nblk);
int build_sell(phg_hip_mat *A);
void free_sell(phg_hip_mat *A);
int ensure_fp32_copy(phg_hip_mat *A);
// block copy for many right-hand sides (block.hip)
int build_bsr(phg_hip_mat *A);
void free_bsr(phg_hip_mat *A);
// dictionary-compressed CSR (dictcsr.hip): shift 8 / 256 slots / blocks of
// DICT_BLOCK_NNZ entries (chunk dictionary) or shift 4 / XS_SLOTS slots /
// blocks of XS_BLOCK_NNZ entries (staged lines)
constexpr int XS_SHIFT = 4, XS_SLOTS = 128, XS_BLOCK_NNZ = 4096;
int build_dict_layout(phg_hip_mat *A, int shift, int max_slots, int block_nnz,
                      DictLayout *L);
int build_dict(phg_hip_mat *A);     // A->cd
int build_xstage(phg_hip_mat *A);   // A->xs
void free_dict_layout(DictLayout *L);
void free_dict(phg_hip_mat *A);     // both
// Re-assembly for new values (assemble.hip): a plan made once per (space,
// problem, eliminated matrix) with device pointers to the kappa, load and
// Dirichlet-value arrays.  It re-allocates the matrix's PHG-order columns and
// values to a capacity bounding every pattern and drops its solver copy
// (q_stale).  refill_enqueue then launches the element pass, row walk, scans,
// compaction and elimination into the matrix's PHG-order arrays, with no
// allocation, synchronisation or memset (capturable); the new entry count is
// at the device address refill_nnz.
struct RefillPlan;
int refill_plan_create(phg_hip_space *sp, const phg_hip_problem *pb, phg_hip_mat *A,
                       const double *d_kappa, const double *d_f, const double *d_g,
                       RefillPlan **out);
int refill_enqueue(const RefillPlan *p, hipStream_t s);
const int64_t *refill_nnz(const RefillPlan *p);
void refill_plan_destroy(RefillPlan *p);
int setup_diagonal(phg_hip_mat *A);
// Rows (or blocks) of one XCD swept together: [0, n) is split into 8 equal
// regions, region k processed by the work-groups the dispatcher places on
// XCD k (blockIdx % 8 == k), work-group j of the XCD taking items
// lo + j, lo + j + w, ... (w work-groups on the XCD).  At any time an XCD
// streams one window of consecutive items (DRAM page locality, as a copy
// kernel), and the x entries a region gathers stay in that XCD's L2.
__device__ inline void xcd_sweep(int64_t n, int64_t &first, int64_t &end,
                                 int64_t &step)
{
    constexpr int64_t kXcd = 8;
    const int64_t nb = gridDim.x, b = blockIdx.x;
    if (nb < kXcd) {
        first = b;
        end = n;
        step = nb;
        return;
    }
    const int64_t x = b % kXcd, j = b / kXcd;
    step = nb / kXcd + (x < nb % kXcd ? 1 : 0);
    first = n * x / kXcd + j;
    end = n * (x + 1) / kXcd;
}
// space-filling curve keys of 21-bit coordinates (solver numbering, element
// order of the matrix-free operator)
__host__ __device__ inline uint64_t curve_spread3(uint64_t v)
{
    // 21 bits -> every third bit of 63
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// Hilbert index of 21-bit coordinates: J. Skilling's transform of the axes
// to the transposed Hilbert integer (AIP Conf. Proc. 707, 381 (2004),
// "Programming the Hilbert curve"), whose bits interleaved give the index.
// Consecutive indices are face neighbours, unlike the Morton curve's jumps.
__host__ __device__ inline uint64_t hilbert3(uint32_t x0, uint32_t x1, uint32_t x2)
{
    uint32_t X[3] = {x0, x1, x2};
    constexpr uint32_t M = 1u << 20;
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (X[i] & Q)
                X[0] ^= P;
            else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q)
            t ^= Q - 1;
#pragma unroll
    for (int i = 0; i < 3; i++)
        X[i] ^= t;
    return curve_spread3(X[0]) << 2 | curve_spread3(X[1]) << 1 | curve_spread3(X[2]);
}

// matrix-free operator (eop.hip; reference data from assemble.hip)
constexpr int EOP_MAX_PART = 2048;   // partial sums one apply writes, at most
constexpr uint32_t EOP_NOT_SHARED = 0xffffffffu;
// T: [7][nb][nb], the six stiffness tensors and the mass tensor
void eop_reference_tensors(int p, double *T);
// the points of the rule those stiffness tensors are summed on: returns
// their number np (at most NPMAX_EOP) with w[np] the weights and
// D[np][nb][3] = d phi_i / d lambda_a - d phi_i / d lambda_0, a = 1..3
constexpr int NPMAX_EOP = 16;
// dwords of one element's index record (the four lane groups' chunks):
// 4 (nj + no + 1) / 2 with nj = (nb + 3) / 4, no = (nb / 2 + 1) / 2, except
// P2, whose lane groups 2 and 3 have one unused x entry and one unused
// result entry each and take two dwords (3 + 3 + 2 + 2)
constexpr int eop_record_dwords(int p)
{
    return p == 2 ? 10 : 4 * ((((p == 1 ? 4 : 20) + 3) / 4 + ((p == 1 ? 4 : 20) / 2 + 1) / 2 + 1) / 2);
}
int eop_rule_gradients(int p, double *w, double *D);
// per element 6 coefficients (mass_coef 0) or 7 (the mass term last)
int eop_coefficients(const phg_hip_space *sp, const double *d_kappa,
                     const int32_t *d_eperm, double mass_coef, float *d_coef,
                     hipStream_t s);
// launches y = A x of the operator in the solver numbering (diag: FP32
// diagonal, for the Dirichlet rows) and writes the eo->grid_b partial sums
// of (x, y) into part_out.  The rows shared between element blocks are set
// by a second kernel when fixup is true; with fixup false they are left to
// the caller, which sums their partials (eo->part) through eo->fxc.  which
// = 1 launches the block kernel only, 2 the fix-up only (timing).  The
// caller checks the launch; the kernels return at once when done is set and
// (*done != 0 or *mode != want)
void eop_apply(const EopData *eo, const float *x, const float *diag, float *y,
               double *part_out, const int32_t *done, const int32_t *mode, int32_t want,
               hipStream_t s, int which = 0, bool fixup = true);
void free_eop(phg_hip_mat *A);
}  // namespace phg
