// Host-side set-up helpers of libftmi (plain C++; called once per mesh or per AMG rebuild).
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ftmi.h"

extern "C" {

// Per-row source lists for ftmi_assemble: row c receives (face << 1 | which, slot) pairs for
// every face coefficient that lands in one of its slots, in face order (the order of the
// reference's sequential loop, flowtrace/native/fvm_kernels.cpp:139-143).
int ftmi_assemble_plan_host(int64_t nc, int64_t nf, const int32_t* indptr, const int32_t* slot_PN,
                            const int32_t* slot_NP, int32_t* src_ptr, int32_t* src) {
    // row of a slot: binary search in indptr
    auto row_of = [&](int32_t slot) {
        int64_t lo = 0, hi = nc;
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) / 2;
            if (indptr[mid] <= slot)
                lo = mid;
            else
                hi = mid;
        }
        return lo;
    };
    std::vector<int32_t> cnt(nc + 1, 0);
    for (int64_t f = 0; f < nf; ++f) {
        cnt[row_of(slot_PN[f]) + 1] += 2;
        cnt[row_of(slot_NP[f]) + 1] += 2;
    }
    for (int64_t c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
    for (int64_t c = 0; c <= nc; ++c) src_ptr[c] = cnt[c];
    std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t r1 = row_of(slot_PN[f]);
        src[pos[r1]++] = (int32_t)(f << 1);
        src[pos[r1]++] = slot_PN[f];
        const int64_t r2 = row_of(slot_NP[f]);
        src[pos[r2]++] = (int32_t)((f << 1) | 1);
        src[pos[r2]++] = slot_NP[f];
    }
    return 0;
}

// Three-pass greedy aggregation (flowtrace/linalg/amg.py:26-57): seeds are nodes whose strong
// neighbourhood is still free; leftovers join the first aggregated strong neighbour; what is
// left becomes a singleton. Sequential by definition: the result depends on the node order.
int64_t ftmi_amg_aggregate_host(int64_t n, const int64_t* ptr, const int64_t* idx, int64_t* agg) {
    int64_t na = 0;
    for (int64_t i = 0; i < n; ++i) agg[i] = -1;
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        bool all_free = true;
        for (int64_t k = ptr[i]; k < ptr[i + 1] && all_free; ++k) all_free = agg[idx[k]] < 0;
        if (!all_free) continue;
        agg[i] = na;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) agg[idx[k]] = na;
        ++na;
    }
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            if (agg[idx[k]] >= 0) {
                agg[i] = agg[idx[k]];
                break;
            }
        }
    }
    for (int64_t i = 0; i < n; ++i)
        if (agg[i] < 0) agg[i] = na++;
    return na;
}

// The same aggregation over a CSR pattern and a strength mask (mask[j] != 0: entry j is a
// strong off-diagonal connection), int32 in and out; size[a] receives each aggregate's size.
// Pass 2 (join the first aggregated strong neighbour, in order), the collection of the nodes
// left for pass 3 and the aggregate sizes share one scan. (Pass 1, the greedy choice of roots,
// is inherently ordered: a block-parallel replay does not re-synchronise with the sequential
// tiling on structured orderings, so it stays one thread.)
int64_t ftmi_amg_aggregate_mask_host(int64_t n, const int32_t* ptr, const int32_t* col, const uint8_t* mask,
                                     int32_t* agg, int32_t* size) {
    if (n <= 0) return 0;
    std::memset(agg, 0xff, (size_t)n * sizeof(int32_t));  // -1
    int32_t na = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        const int32_t e = ptr[i + 1];
        bool all_free = true;
        for (int32_t k = ptr[i]; k < e; ++k) {
            if (mask[k] && agg[col[k]] >= 0) {
                all_free = false;
                break;
            }
        }
        if (!all_free) continue;
        agg[i] = na;
        for (int32_t k = ptr[i]; k < e; ++k)
            if (mask[k]) agg[col[k]] = na;
        size[na] = 0;
        ++na;
    }
    std::vector<int32_t> left;
    for (int64_t i = 0; i < n; ++i) {
        int32_t a = agg[i];
        if (a < 0) {
            for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
                if (mask[k] && agg[col[k]] >= 0) {
                    a = agg[col[k]];
                    break;
                }
            }
            if (a < 0) {
                left.push_back((int32_t)i);
                continue;
            }
            agg[i] = a;
        }
        ++size[a];
    }
    for (const int32_t i : left) {
        agg[i] = na;
        size[na] = 1;
        ++na;
    }
    return na;
}


// ---------------------------------------------------------------- Fluent mesh text sections
// Bulk number parsing for the ASCII sections of a Fluent mesh file (the section structure is
// scanned in Python). Tokens are separated by blanks, tabs and line breaks.

static inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

// Hexadecimal integers (face and cell records); returns the count, or -1 - k when token k is
// not a hexadecimal number, or -2 - cap when there are more than `cap` tokens.
int64_t ftmi_parse_hex(const char* text, int64_t len, int64_t* out, int64_t cap) {
    int64_t k = 0, p = 0;
    while (true) {
        while (p < len && is_blank(text[p])) ++p;
        if (p >= len) break;
        if (k >= cap) return -2 - cap;
        int64_t v = 0;
        const int64_t start = p;
        while (p < len && !is_blank(text[p])) {
            const char c = text[p];
            int d;
            if (c >= '0' && c <= '9')
                d = c - '0';
            else if (c >= 'a' && c <= 'f')
                d = c - 'a' + 10;
            else if (c >= 'A' && c <= 'F')
                d = c - 'A' + 10;
            else
                return -1 - k;
            v = v * 16 + d;
            ++p;
        }
        if (p - start > 15) return -1 - k;  // beyond 60 bits: not a mesh index
        out[k++] = v;
    }
    return k;
}

// Decimal floating-point numbers (node coordinates); same return convention.
int64_t ftmi_parse_f64(const char* text, int64_t len, double* out, int64_t cap) {
    int64_t k = 0, p = 0;
    std::vector<char> tok;
    while (true) {
        while (p < len && is_blank(text[p])) ++p;
        if (p >= len) break;
        if (k >= cap) return -2 - cap;
        const int64_t start = p;
        while (p < len && !is_blank(text[p])) ++p;
        tok.assign(text + start, text + p);
        tok.push_back('\0');
        char* end = nullptr;
        const double v = std::strtod(tok.data(), &end);
        if (end != tok.data() + (p - start)) return -1 - k;
        out[k++] = v;
    }
    return k;
}

}  // extern "C"
