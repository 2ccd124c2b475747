// Dictionary-compressed CSR of the solver matrix.
//
// Rows (in the solver numbering) are grouped into blocks of at most
// block_nnz entries (DICT_BLOCK_NNZ or XS_BLOCK_NNZ) and DICT_BLOCK_ROWS
// rows.  For each block the distinct column groups (column >> shift) its
// entries touch are listed in a dictionary of at most max_slots group bases,
// and every entry stores its column as 16 bits:
// slot << shift | (column & (2^shift - 1)).  Two instances are built:
// 256-column chunks (shift 8, 256 slots), decoded through the dictionary;
// and 16-column lines (shift 4, 128 slots), whose x values the SpMV stages
// in LDS so that the 16-bit code is the LDS index.  In the spatial (Morton)
// numbering a block's columns fall into a few dozen chunks (about 80 lines),
// so the column stream shrinks from 4 to 2 bytes per entry -- the SpMV
// moves 6 instead of 8 bytes per FP32 entry and 10 instead of 12 per FP64
// entry.  Each block's entries start at a multiple of 8 so a work-item loads
// 8 indices and 8 values with 16-byte loads; the padding entries have value
// 0 and index 0 and are counted in the last row of the block (they add +0
// to its sum).
//
// Inside a block, groups of up to 512 entries (one wavefront's share) are
// stored transposed (8 entries per work-item, see the builder); row
// pointers index the untransposed order, which is the solver CSR's storage
// order within every row.  The builder runs on the host over the solver CSR
// once per matrix; a matrix with a row longer than a block or needing more
// than max_slots groups in one row is marked as not representable (the
// autotuner then skips the layout).
#include "phg_hip_internal.h"

#include <algorithm>
#include <vector>

namespace phg {

void free_dict_layout(DictLayout *L)
{
    dfree(L->rowptr);
    dfree(L->blk_row);
    dfree(L->blk_ent);
    dfree(L->blk_dict);
    dfree(L->dict);
    dfree(L->idx);
    dfree(L->val32);
    dfree(L->val64);
    *L = DictLayout();
}

void free_dict(phg_hip_mat *A)
{
    free_dict_layout(&A->cd);
    free_dict_layout(&A->xs);
}

int build_dict(phg_hip_mat *A)
{
    return build_dict_layout(A, 8, 256, DICT_BLOCK_NNZ, &A->cd);
}

int build_xstage(phg_hip_mat *A)
{
    return build_dict_layout(A, XS_SHIFT, XS_SLOTS, XS_BLOCK_NNZ, &A->xs);
}

namespace {

// The host plan of a layout over the solver CSR (rp, ci): block tables,
// dictionary, 16-bit entry codes in their stored (transposed) order, and for
// every stored entry the position of its value in the solver CSR (-1 for
// padding), from which the builder takes the values.  Returns 1, or -1 when
// the matrix is not representable.
struct DictPlan {
    std::vector<int32_t> brow, bent, bdict, dict, crp, src;
    std::vector<uint16_t> idx;
};

int plan_dict(const std::vector<int32_t> &rp, const std::vector<int32_t> &ci, int64_t n,
              int shift, int max_slots, int block_nnz, DictPlan &P)
{
    const int32_t lmask = (1 << shift) - 1;
    const int64_t nnz = rp[(size_t)n];
    const int64_t nchunk = (n + lmask) >> shift;
    std::vector<int32_t> mark((size_t)nchunk, -1), slot((size_t)nchunk, 0);
    std::vector<int32_t> &brow = P.brow, &bent = P.bent, &bdict = P.bdict, &dict = P.dict;
    std::vector<int32_t> &crp = P.crp, &src = P.src;
    std::vector<uint16_t> &idx = P.idx;
    std::vector<int32_t> fresh, tsrc;
    std::vector<uint16_t> tidx;
    crp.assign((size_t)n + 1, 0);
    idx.reserve((size_t)nnz + (size_t)nnz / 64 + 64);
    src.reserve(idx.capacity());
    int64_t r = 0;
    int32_t stamp = 0;
    while (r < n) {
        // a new block
        brow.push_back((int32_t)r);
        bent.push_back((int32_t)idx.size());
        bdict.push_back((int32_t)dict.size());
        const int64_t e0 = (int64_t)idx.size();
        int nslot = 0, nrows = 0;
        while (r < n && nrows < DICT_BLOCK_ROWS) {
            const int64_t j0 = rp[(size_t)r], j1 = rp[(size_t)r + 1];
            const int64_t len = j1 - j0;
            if (len > block_nnz)
                return -1;
            if ((int64_t)idx.size() - e0 + len > block_nnz)
                break;
            // chunks of this row not yet in the block's dictionary
            fresh.clear();
            for (int64_t j = j0; j < j1; j++) {
                int32_t c = ci[(size_t)j] >> shift;
                if (mark[(size_t)c] != stamp && mark[(size_t)c] != -2 - stamp) {
                    mark[(size_t)c] = -2 - stamp;    // tentatively new
                    fresh.push_back(c);
                }
            }
            if (nslot + (int)fresh.size() > max_slots) {
                for (int32_t c : fresh)
                    mark[(size_t)c] = -1;
                if (nrows == 0)
                    return -1;
                break;
            }
            for (int32_t c : fresh) {
                mark[(size_t)c] = stamp;
                slot[(size_t)c] = nslot++;
                dict.push_back(c << shift);
            }
            crp[(size_t)r] = (int32_t)idx.size();
            for (int64_t j = j0; j < j1; j++) {
                int32_t c = ci[(size_t)j];
                idx.push_back((uint16_t)(slot[(size_t)(c >> shift)] << shift | (c & lmask)));
                src.push_back((int32_t)j);
            }
            nrows++;
            r++;
        }
        if (nslot == 0)
            dict.push_back(0);       // padding entries read slot 0
        while (idx.size() % 8) {
            idx.push_back(0);
            src.push_back(-1);
        }
        // store each group of up to 512 entries transposed: memory position
        // 8 l + j holds entry j L + l of the group (L = group size / 8), so
        // the work-item of lane l loads its 8 entries with 16-byte accesses
        // while the j-th gather of a wavefront reads 64 consecutive entries
        {
            const int64_t m = (int64_t)idx.size() - e0;
            for (int64_t g0 = 0; g0 < m; g0 += DICT_GROUP) {
                const int64_t gm = std::min<int64_t>(DICT_GROUP, m - g0), L = gm / 8;
                tidx.assign(idx.begin() + e0 + g0, idx.begin() + e0 + g0 + gm);
                tsrc.assign(src.begin() + e0 + g0, src.begin() + e0 + g0 + gm);
                for (int64_t l = 0; l < L; l++)
                    for (int j = 0; j < 8; j++) {
                        idx[(size_t)(e0 + g0 + 8 * l + j)] = tidx[(size_t)(j * L + l)];
                        src[(size_t)(e0 + g0 + 8 * l + j)] = tsrc[(size_t)(j * L + l)];
                    }
            }
        }
        stamp++;
        if (idx.size() > (size_t)INT32_MAX - block_nnz)
            return -1;
    }
    crp[(size_t)n] = (int32_t)idx.size();
    brow.push_back((int32_t)n);
    bent.push_back((int32_t)idx.size());
    bdict.push_back((int32_t)dict.size());
    return 1;
}

int download_pattern(const phg_hip_mat *A, std::vector<int32_t> &rp, std::vector<int32_t> &ci)
{
    int rc;
    rp.resize((size_t)A->n + 1);
    ci.resize((size_t)A->nnz);
    if ((rc = download(rp.data(), A->q_rowptr, rp.size() * sizeof(int32_t))) ||
        (rc = download(ci.data(), A->q_cols, ci.size() * sizeof(int32_t))))
        return rc;
    return PHG_HIP_OK;
}

}  // namespace

int build_dict_layout(phg_hip_mat *A, int shift, int max_slots, int block_nnz,
                      DictLayout *Lo)
{
    if (Lo->state != 0)
        return PHG_HIP_OK;
    const int64_t nnz = A->nnz;
    int rc;
    std::vector<int32_t> rp, ci;
    std::vector<double> va((size_t)nnz);
    if ((rc = download_pattern(A, rp, ci)) ||
        (rc = download(va.data(), A->q_vals, va.size() * sizeof(double))))
        return rc;
    DictPlan P;
    if (plan_dict(rp, ci, A->n, shift, max_slots, block_nnz, P) < 0) {
        Lo->state = -1;
        return PHG_HIP_OK;
    }
    const int64_t nb = (int64_t)P.brow.size() - 1, ne = (int64_t)P.idx.size();
    std::vector<double> val((size_t)ne);
    std::vector<float> val32((size_t)ne);
    for (size_t i = 0; i < (size_t)ne; i++) {
        val[i] = P.src[i] >= 0 ? va[(size_t)P.src[i]] : 0.0;
        val32[i] = (float)val[i];
    }
    DictLayout L;
    L.shift = shift;
    L.max_slots = max_slots;
    if ((rc = dmalloc((void **)&L.rowptr, P.crp.size() * 4)) ||
        (rc = dmalloc((void **)&L.blk_row, P.brow.size() * 4)) ||
        (rc = dmalloc((void **)&L.blk_ent, P.bent.size() * 4)) ||
        (rc = dmalloc((void **)&L.blk_dict, P.bdict.size() * 4)) ||
        (rc = dmalloc((void **)&L.dict, P.dict.size() * 4)) ||
        (rc = dmalloc((void **)&L.idx, (size_t)ne * 2)) ||
        (rc = dmalloc((void **)&L.val32, (size_t)ne * 4)) ||
        (rc = dmalloc((void **)&L.val64, (size_t)ne * 8)) ||
        (rc = upload(L.rowptr, P.crp.data(), P.crp.size() * 4)) ||
        (rc = upload(L.blk_row, P.brow.data(), P.brow.size() * 4)) ||
        (rc = upload(L.blk_ent, P.bent.data(), P.bent.size() * 4)) ||
        (rc = upload(L.blk_dict, P.bdict.data(), P.bdict.size() * 4)) ||
        (rc = upload(L.dict, P.dict.data(), P.dict.size() * 4)) ||
        (rc = upload(L.idx, P.idx.data(), (size_t)ne * 2)) ||
        (rc = upload(L.val32, val32.data(), (size_t)ne * 4)) ||
        (rc = upload(L.val64, val.data(), (size_t)ne * 8))) {
        free_dict_layout(&L);
        return rc;
    }
    L.nblk = nb;
    L.nent = ne;
    L.ndict = (int64_t)P.dict.size();
    L.state = 1;
    *Lo = L;
    return PHG_HIP_OK;
}

}  // namespace phg
