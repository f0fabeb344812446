/* ORACLE helper (test infrastructure only): the three-pass greedy aggregation of the
 * reference's smoothed-aggregation AMG (flowtrace/linalg/amg.py:26-57), in C so that the
 * CPU baseline can set up AMG hierarchies at a million cells. Loaded by
 * oracle/simplec_oracle.py; the pure-Python loop there is the same algorithm. */
#include <stdint.h>

int64_t oracle_aggregate(int64_t n, const int64_t* ptr, const int64_t* idx, int64_t* agg) {
    int64_t na = 0;
    for (int64_t i = 0; i < n; ++i) agg[i] = -1;
    /* pass 1: a node whose strong neighbours are all free seeds an aggregate with them */
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        int free_nb = 1;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
            if (agg[idx[k]] >= 0) { free_nb = 0; break; }
        if (!free_nb) continue;
        agg[i] = na;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) agg[idx[k]] = na;
        ++na;
    }
    /* pass 2: a leftover joins the aggregate of its first aggregated strong neighbour */
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
            if (agg[idx[k]] >= 0) { agg[i] = agg[idx[k]]; break; }
    }
    /* pass 3: isolated nodes become singletons */
    for (int64_t i = 0; i < n; ++i)
        if (agg[i] < 0) agg[i] = na++;
    return na;
}
