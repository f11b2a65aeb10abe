/*
 * oracle.h -- CPU oracle for the GROUP BY hot path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline. The
 * GPU operator never calls it.
 *
 * Two restatements live here:
 *   orc_groupby_exact  exact GROUP BY: int64/__int128 COUNT/SUM/MIN/MAX,
 *                      Kahan-Neumaier FP64 for DOUBLE SUM/AVG, the reference's
 *                      predicate semantics. The parity target.
 *   orc_port_groupby   the reference's single-column fused CPU path
 *                      (physical_planner.cpp:7613-7931 + direct emit
 *                      :8027-8173): per-thread maps over 122,880-row morsels,
 *                      double accumulators, 4-shard parallel merge. The timed
 *                      CPU baseline.
 *   orc_port_radix_count  RadixCountAgg (radix_count_agg.cpp:17-106) for
 *                      COUNT(*)-only single-key queries.
 */
#ifndef SLOTHDB_ORACLE_H
#define SLOTHDB_ORACLE_H

#include <stdint.h>
#include "../include/slothdb_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_OUT (GBAGG_MAX_GROUP_COLS + GBAGG_MAX_AGGS)

/* Result table. Column c holds 8-byte cells: int64 for INTEGER/BIGINT/VARCHAR
 * (dictionary code) outputs, IEEE double bits for DOUBLE outputs. Rows are in
 * first-appearance order of their group (row order of the input). */
typedef struct {
    uint64_t num_rows;
    uint32_t num_columns;
    int32_t types[ORC_MAX_OUT];
    int64_t *cells[ORC_MAX_OUT];
    uint8_t *valid[ORC_MAX_OUT];
    int32_t overflow; /* 1 if an integer SUM left the int64 range */
} orc_result;

uint64_t orc_hash_u64(uint64_t x);

/* Exact oracle. `threads` >= 1. Returns 0 on success. */
int orc_groupby_exact(const gbagg_query_desc *q, const gbagg_column_view *cols, uint64_t nrows,
                      int threads, orc_result *out);

/* Restatement of the reference's per-thread-map path (single INTEGER/BIGINT
 * group column; COUNT, SUM, AVG, MIN, MAX over INTEGER/BIGINT/DOUBLE).
 * `threads` is the worker count (the reference caps it at 8). When the query
 * carries a TopN hint on the group column or a COUNT aggregate, only the top
 * rows are emitted, through the bounded heap of :8067-8152. */
int orc_port_groupby(const gbagg_query_desc *q, const gbagg_column_view *cols, uint64_t nrows,
                     int threads, orc_result *out);

/* Restatement of RadixCountAgg: COUNT(*) GROUP BY one INTEGER/BIGINT key,
 * 16-way scatter by the hash's top 4 bits, per-radix count maps, optional
 * top-k by count DESC (k <= 0: all groups). Output: keys and counts arrays
 * (malloc'd, caller frees with orc_free). */
int orc_port_radix_count(const int64_t *keys, uint64_t nrows, int threads, int k,
                         int64_t **out_keys, int64_t **out_counts, uint64_t *out_n);

/* Exact top-k of an oracle result by output column `col` (ASC or DESC), ties
 * broken by the group key columns ascending (NULL first). Writes row indices
 * of the result into idx (k entries max); returns the number written. */
uint64_t orc_topn(const orc_result *r, uint32_t num_key_cols, int32_t col, int32_t ascending,
                  uint64_t k, uint64_t *idx);

void orc_result_free(orc_result *r);
void orc_free(void *p);

#ifdef __cplusplus
}
#endif

#endif
