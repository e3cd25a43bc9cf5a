/*
 * CPU restatement of the reference's GPU COO->CSR conversion, used only as a
 * test oracle (never linked into the product library).
 *
 * Follows, step for step:
 *   iterative/cgd/coo_csr.metal:17-46    radix_frequencies  (per 32-key group, 16-bin digit histogram)
 *   iterative/cgd/coo_csr.metal:48-69    vertical_scan      (per-bucket exclusive scan over groups)
 *   iterative/cgd/coo_csr.metal:72-99    scan_histogram     (exclusive scan of the 16 bucket totals)
 *   iterative/cgd/coo_csr.metal:101-142  reorder            (stable scatter; rank among same-digit lanes of the group)
 *   iterative/cgd/coo_csr.metal:144-178  coo_to_csr_compress
 *   iterative/cgd/coo_csr.metal:180-193  row_fixer          (backward fill of empty rows)
 *   iterative/BiCGSTAB/spmv.cpp:62-228   pass schedule: 4-bit digits, a digit is skipped when it lies
 *                                        outside bitlen(num_cols) (low half) / bitlen(num_rows) (high half)
 *
 * Deliberate differences (the reference's behaviour is undefined or buggy there):
 *   - nnz == 0 gives row_ptr all zero (the reference leaves 0xFFFFFFFF, compress never runs);
 *   - the sorted result is taken from wherever the last pass wrote it (the reference swaps it
 *     away when the pass count is odd, spmv.cpp:167-182);
 *   - num_cols == 0 / num_rows == 0 use a zero-bit mask instead of __builtin_clz(0).
 * Returns 0 on success, -1 on allocation failure, -2 when a key lies outside [0,num_rows)x[0,num_cols).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NUM_BUCKETS 16
#define GROUP 32

static uint32_t bitlen32(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

int oracle_lsd_coo_to_csr(const uint64_t *keys, const float *vals, uint64_t nnz,
                          uint32_t num_rows, uint32_t num_cols,
                          uint32_t *row_ptr, uint32_t *col_ind, float *out_vals)
{
    for (uint64_t i = 0; i < nnz; ++i) {
        if ((keys[i] >> 32) >= num_rows || (keys[i] & 0xFFFFFFFFull) >= num_cols) return -2;
    }
    if (nnz == 0) {
        memset(row_ptr, 0, ((size_t)num_rows + 1) * sizeof(uint32_t));
        return 0;
    }
    uint64_t cbits = bitlen32(num_cols), rbits = bitlen32(num_rows);
    uint64_t mask = ((1ull << cbits) - 1ull) | (((1ull << rbits) - 1ull) << 32);

    uint64_t num_groups = (nnz + GROUP - 1) / GROUP;
    uint64_t *k1 = malloc(nnz * sizeof(uint64_t)), *k2 = malloc(nnz * sizeof(uint64_t));
    float *v1 = malloc(nnz * sizeof(float)), *v2 = malloc(nnz * sizeof(float));
    uint32_t *counts = malloc(num_groups * NUM_BUCKETS * sizeof(uint32_t));
    uint32_t *grid_off = malloc(num_groups * NUM_BUCKETS * sizeof(uint32_t));
    if (!k1 || !k2 || !v1 || !v2 || !counts || !grid_off) {
        free(k1); free(k2); free(v1); free(v2); free(counts); free(grid_off);
        return -1;
    }
    memcpy(k1, keys, nnz * sizeof(uint64_t));
    memcpy(v1, vals, nnz * sizeof(float));
    uint64_t *src_k = k1, *dst_k = k2;
    float *src_v = v1, *dst_v = v2;

    for (int shift = 0; shift < 64; shift += 4) {
        if (((0xFull << shift) & mask) == 0) continue;
        /* radix_frequencies */
        memset(counts, 0, num_groups * NUM_BUCKETS * sizeof(uint32_t));
        for (uint64_t i = 0; i < nnz; ++i)
            counts[(i / GROUP) * NUM_BUCKETS + ((src_k[i] >> shift) & 0xF)]++;
        /* vertical_scan */
        uint32_t totals[NUM_BUCKETS];
        for (int b = 0; b < NUM_BUCKETS; ++b) {
            uint32_t sum = 0;
            for (uint64_t g = 0; g < num_groups; ++g) {
                grid_off[g * NUM_BUCKETS + b] = sum;
                sum += counts[g * NUM_BUCKETS + b];
            }
            totals[b] = sum;
        }
        /* scan_histogram */
        uint32_t global_off[NUM_BUCKETS];
        global_off[0] = 0;
        for (int b = 1; b < NUM_BUCKETS; ++b) global_off[b] = global_off[b - 1] + totals[b - 1];
        /* reorder */
        for (uint64_t g = 0; g < num_groups; ++g) {
            uint64_t lo = g * GROUP, hi = lo + GROUP < nnz ? lo + GROUP : nnz;
            for (uint64_t i = lo; i < hi; ++i) {
                uint32_t b = (uint32_t)((src_k[i] >> shift) & 0xF);
                uint32_t rank = 0;
                for (uint64_t j = lo; j < i; ++j) rank += ((src_k[j] >> shift) & 0xF) == b;
                uint64_t o = (uint64_t)global_off[b] + grid_off[g * NUM_BUCKETS + b] + rank;
                dst_k[o] = src_k[i];
                dst_v[o] = src_v[i];
            }
        }
        uint64_t *tk = src_k; src_k = dst_k; dst_k = tk;
        float *tv = src_v; src_v = dst_v; dst_v = tv;
    }
    /* coo_to_csr_compress (row_ptr pre-filled with 0xFFFFFFFF, spmv.cpp:187-201) */
    for (uint64_t r = 0; r <= num_rows; ++r) row_ptr[r] = 0xFFFFFFFFu;
    for (uint64_t i = 0; i < nnz; ++i) {
        uint32_t row = (uint32_t)(src_k[i] >> 32);
        col_ind[i] = (uint32_t)(src_k[i] & 0xFFFFFFFFull);
        out_vals[i] = src_v[i];
        if (i == 0) {
            row_ptr[0] = 0;
            row_ptr[row] = 0;
        } else if (row != (uint32_t)(src_k[i - 1] >> 32)) {
            row_ptr[row] = (uint32_t)i;
        }
    }
    row_ptr[num_rows] = (uint32_t)nnz;
    /* row_fixer */
    for (int64_t i = (int64_t)num_rows - 1; i >= 0; --i)
        if (row_ptr[i] == 0xFFFFFFFFu) row_ptr[i] = row_ptr[i + 1];

    free(k1); free(k2); free(v1); free(v2); free(counts); free(grid_off);
    return 0;
}

/* Sequential FP32 CSR SpMV: y_i = sum over the row in storage order, one float
 * accumulator (the arithmetic of spmv_op, iterative/BiCGSTAB/spmv.metal:11-49,
 * with the 32-lane tree replaced by a left-to-right sum). */
void oracle_spmv_f32_seq(const uint32_t *row_ptr, const uint32_t *col_ind, const float *vals,
                         uint32_t num_rows, const float *x, float *y)
{
    for (uint32_t r = 0; r < num_rows; ++r) {
        float s = 0.0f;
        for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += vals[k] * x[col_ind[k]];
        y[r] = s;
    }
}
