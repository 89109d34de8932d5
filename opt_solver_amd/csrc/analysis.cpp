// Host symbolic analysis producing the AnalysisKeep fields the numeric phase
// reads (solver/include/smf/analysis.hpp:25-144 of the reference): permuted
// lower pattern with an entry map, elimination tree, postorder, column
// counts, fundamental supernodes, relaxed amalgamation and front row sets.
//
// smf's own analysis (solver/src/symbolic_analysis.cpp:434-1291) stays the
// drop-in producer of this data; this implementation is what the library
// uses when no smf analysis is supplied. The ordering is an input (perm,
// new -> old); amalgamation merges a supernode into its parent when it is
// the parent's last child and the merged width stays within nemin (the
// classic relaxed-supernode rule), then children whose contribution block
// costs more to move than to eliminate inside the parent (a cost model of
// its own, in the spirit of smf's, solver/src/supernode_detection.cpp:41,
// 508).
#include "smfgpu_internal.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace smfgpu {

namespace {

// Elimination tree of a lower-CSC pattern (row >= col per entry) by Liu's
// algorithm with path compression over the rows of the upper triangle.
std::vector<int32_t> etree_lower(int32_t n, const std::vector<int32_t> &cp,
                                 const std::vector<int32_t> &ri) {
    // transpose: for each row i, the columns k < i with L(i,k) in pattern
    std::vector<int32_t> rp(n + 1, 0), rc;
    for (int32_t j = 0; j < n; ++j)
        for (int32_t p = cp[j]; p < cp[j + 1]; ++p)
            if (ri[p] > j) ++rp[ri[p] + 1];
    for (int32_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
    rc.resize(rp[n]);
    std::vector<int32_t> fill(rp.begin(), rp.end() - 1);
    for (int32_t j = 0; j < n; ++j)
        for (int32_t p = cp[j]; p < cp[j + 1]; ++p)
            if (ri[p] > j) rc[fill[ri[p]]++] = j;
    std::vector<int32_t> parent(n, -1), anc(n, -1);
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
            int32_t k = rc[p];
            while (k != -1 && k < i) {
                const int32_t next = anc[k];
                anc[k] = i;
                if (next == -1) {
                    parent[k] = i;
                    break;
                }
                k = next;
            }
        }
    }
    return parent;
}

std::vector<int32_t> tree_postorder(int32_t n, const std::vector<int32_t> &parent) {
    std::vector<int32_t> head(n, -1), next(n, -1), post;
    post.reserve(n);
    for (int32_t j = n - 1; j >= 0; --j) {
        if (parent[j] == -1) continue;
        next[j] = head[parent[j]];
        head[parent[j]] = j;
    }
    std::vector<int32_t> stack;
    for (int32_t r = 0; r < n; ++r) {
        if (parent[r] != -1) continue;
        stack.push_back(r);
        while (!stack.empty()) {
            const int32_t p = stack.back();
            const int32_t c = head[p];
            if (c == -1) {
                stack.pop_back();
                post.push_back(p);
            } else {
                head[p] = next[c];
                stack.push_back(c);
            }
        }
    }
    return post;
}

// Column counts |L(:,j)| (diagonal included) of a postordered lower
// pattern (parent[j] > j), by the row-subtree / skeleton method of Gilbert,
// Ng and Peyton (least common ancestors of consecutive leaves found with a
// path-compressed union-find).
std::vector<int64_t> column_counts(int32_t n, const std::vector<int32_t> &cp,
                                   const std::vector<int32_t> &ri,
                                   const std::vector<int32_t> &parent) {
    std::vector<int64_t> delta(n, 0);
    std::vector<int32_t> first(n, -1), maxfirst(n, -1), prevleaf(n, -1), anc(n);
    // postorder is the identity here
    for (int32_t k = 0; k < n; ++k) {
        int32_t j = k;
        delta[j] = (first[j] == -1) ? 1 : 0;
        for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
    }
    std::iota(anc.begin(), anc.end(), 0);
    for (int32_t k = 0; k < n; ++k) {
        const int32_t j = k;
        if (parent[j] != -1) --delta[parent[j]];
        for (int32_t p = cp[j]; p < cp[j + 1]; ++p) {
            const int32_t i = ri[p];
            if (i <= j || first[j] <= maxfirst[i]) continue;  // j not a leaf of row subtree i
            maxfirst[i] = first[j];
            const int32_t jprev = prevleaf[i];
            prevleaf[i] = j;
            ++delta[j];
            if (jprev == -1) continue;  // first leaf: the path to i counts once
            int32_t q = jprev;
            while (q != anc[q]) q = anc[q];
            for (int32_t s = jprev; s != q;) {
                const int32_t sp = anc[s];
                anc[s] = q;
                s = sp;
            }
            --delta[q];
        }
        if (parent[j] != -1) anc[j] = parent[j];
    }
    for (int32_t j = 0; j < n; ++j)
        if (parent[j] != -1) delta[parent[j]] += delta[j];
    return delta;
}

// Supernodes of the column ranges [starts[k], starts[k+1]) (a topological
// order of the elimination tree `parent`): parents, children (descending
// order) and front rows (pivot columns, then the rows below from the
// pattern and the children's fronts, sorted).
int build_tree(const std::vector<int32_t> &starts, const std::vector<int32_t> &parent, Analysis &out) {
    const int32_t n = out.n;
    const int32_t nsn = static_cast<int32_t>(starts.size()) - 1;
    out.sn_start.assign(starts.begin(), starts.end() - 1);
    out.sn_end.assign(starts.begin() + 1, starts.end());
    std::vector<int32_t> col2sn(n);
    for (int32_t k = 0; k < nsn; ++k)
        for (int32_t j = out.sn_start[k]; j < out.sn_end[k]; ++j) col2sn[j] = k;
    out.sn_parent.assign(nsn, -1);
    for (int32_t k = 0; k < nsn; ++k) {
        const int32_t pj = parent[out.sn_end[k] - 1];
        out.sn_parent[k] = pj < 0 ? -1 : col2sn[pj];
        if (out.sn_parent[k] >= 0 && out.sn_parent[k] <= k)
            return set_error(SMFGPU_ERR_INVALID_ARG, "analyse: supernode order is not topological");
    }
    out.child_ptr.assign(nsn + 1, 0);
    for (int32_t k = 0; k < nsn; ++k)
        if (out.sn_parent[k] >= 0) ++out.child_ptr[out.sn_parent[k] + 1];
    for (int32_t k = 0; k < nsn; ++k) out.child_ptr[k + 1] += out.child_ptr[k];
    out.child_idx.assign(out.child_ptr[nsn], 0);
    {
        std::vector<int32_t> fill(out.child_ptr.begin(), out.child_ptr.end() - 1);
        for (int32_t k = nsn - 1; k >= 0; --k)
            if (out.sn_parent[k] >= 0) out.child_idx[fill[out.sn_parent[k]]++] = k;
    }
    out.front_row_ptr.assign(nsn + 1, 0);
    out.front_row_idx.clear();
    std::vector<int32_t> mark(n, -1);
    std::vector<int32_t> below;
    for (int32_t k = 0; k < nsn; ++k) {
        const int32_t a = out.sn_start[k], b = out.sn_end[k];
        below.clear();
        for (int32_t j = a; j < b; ++j)
            for (int32_t p = out.pcol_ptr[j]; p < out.pcol_ptr[j + 1]; ++p) {
                const int32_t r = out.prow_idx[p];
                if (r >= b && mark[r] != k) {
                    mark[r] = k;
                    below.push_back(r);
                }
            }
        for (int32_t c = out.child_ptr[k]; c < out.child_ptr[k + 1]; ++c) {
            const int32_t ch = out.child_idx[c];
            for (int32_t t = out.front_row_ptr[ch] + (out.sn_end[ch] - out.sn_start[ch]);
                 t < out.front_row_ptr[ch + 1]; ++t) {
                const int32_t r = out.front_row_idx[t];
                if (r >= b && mark[r] != k) {
                    mark[r] = k;
                    below.push_back(r);
                }
            }
        }
        std::sort(below.begin(), below.end());
        for (int32_t j = a; j < b; ++j) out.front_row_idx.push_back(j);
        out.front_row_idx.insert(out.front_row_idx.end(), below.begin(), below.end());
        if (out.front_row_idx.size() > static_cast<size_t>(INT32_MAX))
            return set_error(SMFGPU_ERR_ALLOC, "analyse: front rows exceed 2^31");
        out.front_row_ptr[k + 1] = static_cast<int32_t>(out.front_row_idx.size());
    }
    out.nsup = nsn;
    return SMFGPU_OK;
}

// Cost-model amalgamation: child c joins its parent s when eliminating c's
// p_c columns inside s's front costs less than moving c's contribution
// block. The merged front has f_s + p_c rows (c's contribution rows are
// rows of s), so the extra elimination work is p_c ((f_s + p_c)^2 - f_c^2)
// flops, against about 12 q_c^2 bytes (q_c = f_c - p_c) of the block's
// update, store and extend-add; with Schur updates near 30 TFLOP/s and
// those moves near 3 TB/s, merge when the flops stay below 100 q_c^2. It
// removes chains of narrow separators whose fronts carry the same large
// row set (KKT systems whose constraint rows couple the whole grid).
// Only blocks of at least 1024 rows are considered (below that the per-front
// costs the model leaves out dominate, and merging slowed 3D grids down).
// Bottom-up, a merged child's own merges count in its sizes; not applied
// with nemin = 1 (exact fundamental supernodes). On a merge,
// out.perm becomes the postorder of the merged tree (each group's
// supernodes in their order) and starts the groups' column ranges.
bool amalgamate(Analysis &out, std::vector<int32_t> &starts) {
    constexpr double kMergeRatio = 100.0;
    constexpr double kMergeMinCb = 1024.0;  // smaller blocks: per-front costs the model leaves out dominate
    const int32_t nsn = out.nsup;
    std::vector<double> fz(nsn), pz(nsn);
    for (int32_t k = 0; k < nsn; ++k) {
        fz[k] = out.front_row_ptr[k + 1] - out.front_row_ptr[k];
        pz[k] = out.sn_end[k] - out.sn_start[k];
    }
    std::vector<int32_t> grp(nsn);
    std::iota(grp.begin(), grp.end(), 0);
    bool any = false;
    for (int32_t k = 0; k < nsn; ++k)
        for (int32_t t = out.child_ptr[k]; t < out.child_ptr[k + 1]; ++t) {
            const int32_t c = out.child_idx[t];
            const double qc = fz[c] - pz[c];
            const double fm = fz[k] + pz[c];
            const double extra = pz[c] * (fm * fm - fz[c] * fz[c]);
            if (qc >= kMergeMinCb && extra < kMergeRatio * qc * qc) {
                grp[c] = k;
                fz[k] = fm;
                pz[k] += pz[c];
                any = true;
            }
        }
    if (!any) return false;
    // group roots, members (ascending) and the merged tree
    std::vector<int32_t> root(nsn);
    for (int32_t k = nsn - 1; k >= 0; --k) root[k] = grp[k] == k ? k : root[grp[k]];
    std::vector<std::vector<int32_t>> members(nsn), kids(nsn);
    for (int32_t k = 0; k < nsn; ++k) members[root[k]].push_back(k);
    std::vector<int32_t> tops;
    for (int32_t k = 0; k < nsn; ++k) {
        if (root[k] != k) continue;
        if (out.sn_parent[k] < 0) tops.push_back(k);
        else kids[root[out.sn_parent[k]]].push_back(k);
    }
    // postorder of the merged tree (children in ascending order)
    std::vector<int32_t> newperm, nstarts;
    newperm.reserve(out.n);
    std::vector<std::pair<int32_t, size_t>> stack;
    for (int32_t t : tops) {
        stack.emplace_back(t, 0);
        while (!stack.empty()) {
            auto &top = stack.back();
            if (top.second < kids[top.first].size()) {
                const int32_t ch = kids[top.first][top.second++];
                stack.emplace_back(ch, 0);
                continue;
            }
            nstarts.push_back(static_cast<int32_t>(newperm.size()));
            for (int32_t m : members[top.first])
                for (int32_t j = out.sn_start[m]; j < out.sn_end[m]; ++j) newperm.push_back(out.perm[j]);
            stack.pop_back();
        }
    }
    nstarts.push_back(out.n);
    out.perm = newperm;
    starts = nstarts;
    return true;
}

}  // namespace

int analyse(int32_t n, const int32_t *col_ptr, const int32_t *row_idx, const int32_t *perm_in,
            int32_t nemin, Analysis &out) {
    if (n < 0 || (n > 0 && (col_ptr == nullptr || row_idx == nullptr)))
        return set_error(SMFGPU_ERR_INVALID_ARG, "analyse: bad matrix");
    if (nemin <= 0) nemin = 32;
    const int64_t nnz = n > 0 ? col_ptr[n] : 0;
    std::vector<int32_t> perm(n), iperm(n, -1);
    if (perm_in) {
        for (int32_t i = 0; i < n; ++i) {
            const int32_t o = perm_in[i];
            if (o < 0 || o >= n || iperm[o] != -1)
                return set_error(SMFGPU_ERR_INVALID_ARG, "analyse: perm is not a permutation");
            perm[i] = o;
            iperm[o] = i;
        }
    } else {
        std::iota(perm.begin(), perm.end(), 0);
        std::iota(iperm.begin(), iperm.end(), 0);
    }
    // Pattern of P A P^T (lower), as (new col, new row) pairs.
    auto permuted_pattern = [&](const std::vector<int32_t> &ip, std::vector<int32_t> &cp,
                                std::vector<int32_t> &ri, std::vector<int32_t> *slot) {
        cp.assign(n + 1, 0);
        std::vector<int32_t> nc(nnz), nr(nnz);
        for (int32_t j = 0; j < n; ++j)
            for (int64_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p) {
                const int32_t i = row_idx[p];
                if (i < 0 || i >= n) {
                    nc[p] = -1;
                    continue;
                }
                const int32_t a = ip[i], b = ip[j];
                nc[p] = std::min(a, b);
                nr[p] = std::max(a, b);
                ++cp[nc[p] + 1];
            }
        for (int32_t j = 0; j < n; ++j) cp[j + 1] += cp[j];
        // counting sort by column, then sort rows inside each column and
        // merge duplicates
        std::vector<int64_t> order(cp[n]);
        std::vector<int32_t> fill(cp.begin(), cp.end() - 1);
        for (int64_t p = 0; p < nnz; ++p)
            if (nc[p] >= 0) order[fill[nc[p]]++] = p;
        std::vector<int32_t> ncp(n + 1, 0);
        ri.clear();
        ri.reserve(cp[n]);
        if (slot) slot->assign(nnz, -1);
        for (int32_t j = 0; j < n; ++j) {
            auto b = order.begin() + cp[j], e = order.begin() + cp[j + 1];
            std::sort(b, e, [&](int64_t x, int64_t y) { return nr[x] < nr[y] || (nr[x] == nr[y] && x < y); });
            const size_t colstart = ri.size();
            for (auto it = b; it != e; ++it) {
                const int32_t r = nr[*it];
                if (ri.size() == colstart || ri.back() != r) ri.push_back(r);
                if (slot) (*slot)[*it] = static_cast<int32_t>(ri.size()) - 1;
            }
            ncp[j + 1] = static_cast<int32_t>(ri.size());
        }
        cp = ncp;
    };
    std::vector<int32_t> cp0, ri0;
    permuted_pattern(iperm, cp0, ri0, nullptr);
    std::vector<int32_t> parent0 = etree_lower(n, cp0, ri0);
    std::vector<int32_t> post = tree_postorder(n, parent0);
    // compose the postorder into the ordering
    std::vector<int32_t> perm2(n), iperm2(n);
    for (int32_t k = 0; k < n; ++k) perm2[k] = perm[post[k]];
    for (int32_t k = 0; k < n; ++k) iperm2[perm2[k]] = k;
    out.perm = perm2;
    permuted_pattern(iperm2, out.pcol_ptr, out.prow_idx, &out.orig_to_slot);
    std::vector<int32_t> parent = etree_lower(n, out.pcol_ptr, out.prow_idx);
    std::vector<int64_t> cc = column_counts(n, out.pcol_ptr, out.prow_idx, parent);

    // fundamental supernodes
    std::vector<int32_t> nchild(n, 0);
    for (int32_t j = 0; j < n; ++j)
        if (parent[j] != -1) ++nchild[parent[j]];
    std::vector<int32_t> sstart;
    for (int32_t j = 0; j < n; ++j) {
        const bool cont = j > 0 && parent[j - 1] == j && nchild[j] == 1 && cc[j - 1] == cc[j] + 1;
        if (!cont) sstart.push_back(j);
    }
    sstart.push_back(n);
    int32_t ns = static_cast<int32_t>(sstart.size()) - 1;
    std::vector<int32_t> col2sn(n);
    for (int32_t s = 0; s < ns; ++s)
        for (int32_t j = sstart[s]; j < sstart[s + 1]; ++j) col2sn[j] = s;
    std::vector<int32_t> sparent(ns, -1);
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t pj = parent[sstart[s + 1] - 1];
        sparent[s] = pj < 0 ? -1 : col2sn[pj];
    }
    // relaxed amalgamation: merge s into its parent p when s ends where p
    // starts (contiguous columns) and the merged width stays within nemin.
    std::vector<int32_t> merged_into(ns, -1);
    std::vector<int32_t> width(ns), send(ns);
    for (int32_t s = 0; s < ns; ++s) {
        width[s] = sstart[s + 1] - sstart[s];
        send[s] = sstart[s + 1];
    }
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t p = sparent[s];
        if (p < 0 || send[s] != sstart[p]) continue;
        if (width[s] + width[p] <= nemin) {
            sstart[p] = sstart[s];
            width[p] += width[s];
            merged_into[s] = p;
        }
    }
    // renumber surviving supernodes (in column order = postorder)
    std::vector<int32_t> order_s;
    for (int32_t s = 0; s < ns; ++s)
        if (merged_into[s] == -1) order_s.push_back(s);
    std::sort(order_s.begin(), order_s.end(), [&](int32_t a, int32_t b) { return sstart[a] < sstart[b]; });
    std::vector<int32_t> starts;
    for (int32_t s : order_s) starts.push_back(sstart[s]);
    starts.push_back(n);
    out.n = n;
    int rc;
    if ((rc = build_tree(starts, parent, out))) return rc;
    if (nemin > 1 && amalgamate(out, starts)) {
        // columns renumbered so that every merged group is contiguous: the
        // same elimination tree in another topological order, so the same
        // fill
        std::vector<int32_t> ip(n);
        for (int32_t k = 0; k < n; ++k) ip[out.perm[k]] = k;
        permuted_pattern(ip, out.pcol_ptr, out.prow_idx, &out.orig_to_slot);
        parent = etree_lower(n, out.pcol_ptr, out.prow_idx);
        if ((rc = build_tree(starts, parent, out))) return rc;
    }
    return SMFGPU_OK;
}


void Analysis::describe(smfgpu_analysis_desc &d) const {
    d.n = n;
    d.nsup = nsup;
    d.sn_start = sn_start.data();
    d.sn_end = sn_end.data();
    d.sn_parent = sn_parent.data();
    d.child_ptr = child_ptr.data();
    d.child_idx = child_idx.data();
    d.front_row_ptr = front_row_ptr.data();
    d.front_row_idx = front_row_idx.data();
    d.pcol_ptr = pcol_ptr.data();
    d.prow_idx = prow_idx.data();
    d.n_orig = static_cast<int64_t>(orig_to_slot.size());
    d.orig_to_slot = orig_to_slot.data();
    d.perm = perm.data();
}

}  // namespace smfgpu
