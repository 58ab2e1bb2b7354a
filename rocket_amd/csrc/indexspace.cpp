// Pair enumeration over the upper triangle of the M x M pair matrix: Rocket's
// index spaces (R/indexspace/HilbertCurveIndexSpace.java,
// R/indexspace/TilingIndexSpace.java) and pair counting (R/util/Util.java:51-77),
// as host C++ behind the C ABI.  One divide() step is exposed exactly as the
// reference defines it (children to spawn plus entries emitted directly), and
// a driver enumerates every pair in depth-first (curve) order.
#include <cstdint>
#include <vector>

#include "../../include/prnu_api.h"

namespace prnu {
void set_error(const char *fmt, ...);
}

namespace {

// HilbertCurveIndexSpace.java:10-20: {p, q, next rule} for the 4 quadrants.
const int kRules[4][4][3] = {
    {{0, 0, 3}, {0, 1, 0}, {1, 1, 0}, {1, 0, 1}},  // A
    {{1, 1, 2}, {0, 1, 1}, {0, 0, 1}, {1, 0, 0}},  // B
    {{1, 1, 1}, {1, 0, 2}, {0, 0, 2}, {0, 1, 3}},  // C
    {{0, 0, 0}, {1, 0, 3}, {1, 1, 3}, {0, 1, 2}},  // D
};

int64_t tri(int64_t n) { return n * (n + 1) / 2; }

int64_t tri_rect(int64_t rows, int64_t cols) {
    if (cols > rows) return tri(rows) + rows * (cols - rows);
    return tri(cols);
}

int64_t tri_range(int64_t rb, int64_t re, int64_t cb, int64_t ce, bool diag) {
    if (!diag) return tri_range(rb + 1, re + 1, cb, ce, true);
    if (rb >= re && cb >= ce) return 0;
    return tri_rect(rb, cb) + tri_rect(re, ce) - tri_rect(rb, ce) - tri_rect(re, cb);
}

struct Sink {
    std::vector<prnu_index_node> *children;  // null: recurse into children instead
    int32_t *pairs;
    int64_t cap;
    int64_t count;
    void entry(int32_t i, int32_t j) {
        if (count < cap) {
            pairs[2 * count] = i;
            pairs[2 * count + 1] = j;
        }
        ++count;
    }
};

int64_t node_size(const prnu_index_node &n, bool diag) {
    return tri_range(n.left_offset, n.left_offset + n.left_length, n.right_offset, n.right_offset + n.right_length,
                     diag);
}

void hilbert_divide(const prnu_index_node &n, int64_t min_split, bool diag, Sink &out);

// HilbertCurveIndexSpace.launchSplits (:51-98)
void hilbert_splits(const prnu_index_node &n, int64_t min_split, bool diag, Sink &out) {
    const int lh = n.left_length / 2, rh = n.right_length / 2;
    for (int k = 0; k < 4; ++k) {
        const int p = kRules[n.rule][k][0], q = kRules[n.rule][k][1], next = kRules[n.rule][k][2];
        const int l0 = n.left_offset + (p == 0 ? 0 : (lh < n.left_length ? lh : n.left_length));
        const int l1 = n.left_offset + (p == 0 ? lh : n.left_length);
        const int r0 = n.right_offset + (q == 0 ? 0 : (rh < n.right_length ? rh : n.right_length));
        const int r1 = n.right_offset + (q == 0 ? rh : n.right_length);
        const int ln = l1 - l0, rn = r1 - r0;
        if (ln == 0 || rn == 0) continue;
        prnu_index_node c{l0, ln, r0, rn, next};
        if (node_size(c, diag) == 0) continue;
        if ((int64_t)ln * rn < min_split || !out.children)
            hilbert_divide(c, min_split, diag, out);
        else
            out.children->push_back(c);
    }
}

// HilbertCurveIndexSpace.divide / launchEntry (:100-113)
void hilbert_divide(const prnu_index_node &n, int64_t min_split, bool diag, Sink &out) {
    if (n.left_length > 1 || n.right_length > 1) {
        hilbert_splits(n, min_split, diag, out);
    } else if (n.left_offset < n.right_offset || (n.left_offset == n.right_offset && diag)) {
        out.entry(n.left_offset, n.right_offset);
    }
}

// TilingIndexSpace.divide (:55-85); node = rows [lo, lo+ll) x cols [ro, ro+rl)
void tiling_divide(const prnu_index_node &n, int tile, bool diag, Sink &out) {
    const int rb = n.left_offset, re = n.left_offset + n.left_length;
    const int cb = n.right_offset, ce = n.right_offset + n.right_length;
    if (n.left_length > tile || n.right_length > tile) {
        for (int i = rb; i < re; i += tile)
            for (int j = cb; j < ce; j += tile) {
                prnu_index_node c{i, (i + tile < re ? i + tile : re) - i, j, (j + tile < ce ? j + tile : ce) - j, 0};
                if (node_size(c, diag) == 0) continue;
                if (out.children)
                    out.children->push_back(c);
                else
                    tiling_divide(c, tile, diag, out);
            }
    } else {
        for (int i = rb; i < re; ++i)
            for (int j = cb; j < ce; ++j)
                if (i < j || (i == j && diag)) out.entry(i, j);
    }
}

}  // namespace

extern "C" {

int64_t prnu_triangle_size(int row_begin, int row_end, int col_begin, int col_end, int include_diagonal) {
    return tri_range(row_begin, row_end, col_begin, col_end, include_diagonal != 0);
}

int prnu_index_divide(int kind, const prnu_index_node *node, int param, int include_diagonal,
                      prnu_index_node *children, int children_cap, int *n_children, int32_t *pairs,
                      int64_t pairs_cap, int64_t *n_pairs) {
    if (!node || !n_children || !n_pairs || (kind != PRNU_INDEX_HILBERT && kind != PRNU_INDEX_TILING) ||
        param < 1 || node->left_length < 0 || node->right_length < 0 || node->rule < 0 || node->rule > 3) {
        prnu::set_error("index_divide: invalid arguments");
        return PRNU_EINVAL;
    }
    std::vector<prnu_index_node> kids;
    Sink s{&kids, pairs, pairs ? pairs_cap : 0, 0};
    if (kind == PRNU_INDEX_HILBERT)
        hilbert_divide(*node, param, include_diagonal != 0, s);
    else
        tiling_divide(*node, param, include_diagonal != 0, s);
    *n_children = (int)kids.size();
    *n_pairs = s.count;
    if ((int)kids.size() > children_cap || s.count > s.cap) {
        prnu::set_error("index_divide: %d children / %lld pairs exceed the capacity given", (int)kids.size(),
                        (long long)s.count);
        return PRNU_EINVAL;
    }
    for (size_t k = 0; k < kids.size(); ++k) children[k] = kids[k];
    return PRNU_OK;
}

int prnu_index_pairs(int kind, int count, int param, int include_diagonal, int32_t *pairs, int64_t pairs_cap,
                     int64_t *n_pairs) {
    if (!n_pairs || count < 0 || param < 1 || (kind != PRNU_INDEX_HILBERT && kind != PRNU_INDEX_TILING)) {
        prnu::set_error("index_pairs: invalid arguments");
        return PRNU_EINVAL;
    }
    prnu_index_node root{0, count, 0, count, 0};
    Sink s{nullptr, pairs, pairs ? pairs_cap : 0, 0};
    if (count > 0) {
        if (kind == PRNU_INDEX_HILBERT)
            hilbert_divide(root, param, include_diagonal != 0, s);
        else
            tiling_divide(root, param, include_diagonal != 0, s);
    }
    *n_pairs = s.count;
    if (pairs && s.count > pairs_cap) {
        prnu::set_error("index_pairs: %lld pairs exceed the capacity %lld", (long long)s.count, (long long)pairs_cap);
        return PRNU_EINVAL;
    }
    return PRNU_OK;
}

}  // extern "C"
