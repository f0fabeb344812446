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
// strong off-diagonal connection; MASK = false: every entry is strong, the graph given as its
// own CSR), int32 in and out; size[a] receives each aggregate's size. Pass 2 (join the first
// aggregated strong neighbour, in order), the collection of the nodes left for pass 3 and the
// aggregate sizes share one scan. (Pass 1, the greedy choice of roots, is inherently ordered:
// a change of one strength entry can move the roots of a large part of the rows after it, so
// neither a block-parallel replay nor an incremental one re-synchronises with the sequential
// result; it stays one thread.)
}  // extern "C"

template <bool MASK>
static int64_t aggregate_rows(int64_t n, const int32_t* ptr, const int32_t* col, const uint8_t* mask, int32_t* agg,
                              int32_t* size) {
    if (n <= 0) return 0;
    std::memset(agg, 0xff, (size_t)n * sizeof(int32_t));  // -1
    int32_t na = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        const int32_t e = ptr[i + 1];
        bool all_free = true;
        for (int32_t k = ptr[i]; k < e; ++k) {
            if ((!MASK || mask[k]) && agg[col[k]] >= 0) {
                all_free = false;
                break;
            }
        }
        if (!all_free) continue;
        agg[i] = na;
        for (int32_t k = ptr[i]; k < e; ++k)
            if (!MASK || mask[k]) agg[col[k]] = na;
        size[na] = 0;
        ++na;
    }
    std::vector<int32_t> left;
    for (int64_t i = 0; i < n; ++i) {
        int32_t a = agg[i];
        if (a < 0) {
            for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
                if ((!MASK || mask[k]) && agg[col[k]] >= 0) {
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

extern "C" {

int64_t ftmi_amg_aggregate_mask_host(int64_t n, const int32_t* ptr, const int32_t* col, const uint8_t* mask,
                                     int32_t* agg, int32_t* size) {
    return aggregate_rows<true>(n, ptr, col, mask, agg, size);
}

int64_t ftmi_amg_aggregate_csr_host(int64_t n, const int32_t* s_ptr, const int32_t* s_col, int32_t* agg,
                                    int32_t* size) {
    return aggregate_rows<false>(n, s_ptr, s_col, nullptr, agg, size);
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

// ------------------------------------------------------------------ Fluent section index
// One pass over a Fluent mesh file that finds every top-level section and the byte spans a
// decoder needs, so the per-section decoding can work on spans without re-scanning.
//
// Table row (kFluentRow int64 per section): index, header begin, header end, body begin,
// body end, flags. The header is the text inside the first inner group of a grid section
// ("(zone first last type ...)"); the body is the data group: its text for ASCII sections,
// and for binary sections (index >= 2000) the bytes from after its '(' up to the
// "End of Binary Section" marker (the decoder takes the length the header implies). For the
// dimension section the header is its number; for zone-name sections the first inner group.
// Sections that are neither (comments, case settings, unknown indices) are skipped: ASCII ones
// as balanced groups whose quoted strings may hold parentheses, binary ones up to their marker.
// Bytes outside any section (a stray line) are skipped up to the next newline.
namespace {

constexpr int kFluentRow = 6;
constexpr int kHasBody = 1;

struct FluentScan {
    const char* b;
    int64_t n, p = 0;
    int64_t err = 0;  // first error position + 1 (0: none)
    int code = 0;     // what went wrong

    bool fail(int c) {
        if (!err) {
            err = p + 1;
            code = c;
        }
        return false;
    }
    void blanks() {
        while (p < n && is_blank(b[p])) ++p;
    }
    bool take(char c) {
        blanks();
        if (p >= n || b[p] != c) return fail(1);
        ++p;
        return true;
    }
    // a word: up to a blank or a parenthesis
    void word(int64_t& w0, int64_t& w1) {
        blanks();
        w0 = p;
        while (p < n && !is_blank(b[p]) && b[p] != '(' && b[p] != ')') ++p;
        w1 = p;
    }
    // "( ... )" with no inner parentheses: the text span, cursor after ')'
    bool flat_group(int64_t& g0, int64_t& g1) {
        if (!take('(')) return false;
        g0 = p;
        while (p < n && b[p] != ')') {
            if (b[p] == '(') return fail(2);
            ++p;
        }
        if (p >= n) return fail(3);
        g1 = p++;
        return true;
    }
    // skip from just after an opening '(' to after its matching ')', strings honoured
    bool balanced_rest() {
        int depth = 1;
        while (depth && p < n) {
            const char c = b[p];
            if (c == '"') {
                int64_t q = p + 1;
                while (q < n && (b[q] != '"' || b[q - 1] == '\\')) ++q;
                if (q >= n) return fail(4);
                p = q + 1;
                continue;
            }
            depth += (c == '(') - (c == ')');
            ++p;
        }
        return depth == 0 ? true : fail(3);
    }
    // find the binary end marker from p; marker position in m, cursor after "... <index>)"
    bool binary_end(int64_t& m) {
        static const char kMark[] = "End of Binary Section";
        const int64_t L = (int64_t)sizeof(kMark) - 1;
        for (int64_t q = p; q + L <= n; ++q) {
            if (b[q] == 'E' && std::memcmp(b + q, kMark, (size_t)L) == 0) {
                m = q;
                p = q + L;
                int64_t w0, w1;
                word(w0, w1);
                return take(')');
            }
        }
        return fail(5);
    }
};

bool is_grid_section(int64_t idx) {
    switch (idx) {
    case 10: case 2010: case 3010: case 12: case 2012: case 3012:
    case 13: case 2013: case 3013: case 18: case 2018:
        return true;
    default:
        return false;
    }
}

}  // namespace

extern "C" int64_t ftmi_fluent_index(const char* buf, int64_t len, int64_t* table, int64_t cap, int64_t* err_pos) {
    FluentScan s{buf, len};
    int64_t rows = 0;
    auto emit = [&](int64_t idx, int64_t h0, int64_t h1, int64_t b0, int64_t b1, int64_t flags) {
        if (rows < cap) {
            int64_t* r = table + rows * kFluentRow;
            r[0] = idx, r[1] = h0, r[2] = h1, r[3] = b0, r[4] = b1, r[5] = flags;
        }
        ++rows;
    };
    bool ok = true;
    while (ok) {
        s.blanks();
        if (s.p >= s.n) break;
        if (buf[s.p] != '(') {  // stray text between sections: skip the line
            while (s.p < s.n && buf[s.p] != '\n') ++s.p;
            continue;
        }
        ++s.p;
        int64_t w0, w1;
        s.word(w0, w1);
        int64_t idx = 0;
        bool numeric = w1 > w0 && w1 - w0 < 10;
        for (int64_t q = w0; numeric && q < w1; ++q) {
            if (buf[q] < '0' || buf[q] > '9') numeric = false;
            else idx = idx * 10 + (buf[q] - '0');
        }
        if (!numeric) {
            ok = s.balanced_rest();
            continue;
        }
        if (idx == 2) {  // (2 dim)
            int64_t d0, d1;
            s.word(d0, d1);
            ok = s.take(')');
            if (ok) emit(idx, d0, d1, 0, 0, 0);
        } else if (is_grid_section(idx)) {
            int64_t h0, h1;
            ok = s.flat_group(h0, h1);
            if (!ok) break;
            s.blanks();
            if (s.p < s.n && buf[s.p] == ')') {  // a declaration without data
                ++s.p;
                emit(idx, h0, h1, 0, 0, 0);
                continue;
            }
            if (idx < 2000) {
                int64_t g0, g1;
                ok = s.flat_group(g0, g1) && s.take(')');
                if (ok) emit(idx, h0, h1, g0, g1, kHasBody);
            } else {
                ok = s.take('(');
                const int64_t b0 = s.p;
                int64_t m = 0;
                ok = ok && s.binary_end(m);
                if (ok) emit(idx, h0, h1, b0, m, kHasBody);
            }
        } else if (idx == 39 || idx == 45) {  // (45 (id type name)(...))
            int64_t h0, h1;
            ok = s.flat_group(h0, h1);
            while (ok) {
                s.blanks();
                if (s.p < s.n && buf[s.p] == '(') {
                    ++s.p;
                    ok = s.balanced_rest();
                } else {
                    break;
                }
            }
            ok = ok && s.take(')');
            if (ok) emit(idx, h0, h1, 0, 0, 0);
        } else if (idx >= 2000) {
            int64_t m = 0;
            ok = s.binary_end(m);
        } else {
            ok = s.balanced_rest();
        }
    }
    if (!ok || s.err) {
        *err_pos = s.err ? s.err - 1 : s.p;
        return -(int64_t)(s.code ? s.code : 9);
    }
    *err_pos = -1;
    return rows > cap ? -100 - rows : rows;
}

// Start of every record of a mixed (polygonal) face section: record i begins with its node
// count k_i and holds k_i + 3 integers (count, nodes, c0, c1). Returns the integers used, or -1
// when the records run past nvals.
extern "C" int64_t ftmi_fluent_walk(const int64_t* vals, int64_t nvals, int64_t nrec, int64_t* starts) {
    int64_t p = 0;
    for (int64_t i = 0; i < nrec; ++i) {
        if (p >= nvals || vals[p] < 0) return -1;
        starts[i] = p;
        p += vals[p] + 3;
    }
    return p <= nvals ? p : -1;
}
