// Multicolour ILU(0) ordering (host): greedy colouring of the symmetrised pattern graph in
// smallest-last order (Matula-Beck degeneracy order, reversed; Saad, Iterative Methods for
// Sparse Linear Systems, 2nd ed., section 12.4 multicolour orderings), rows sorted by (colour,
// Morton code of their vertex), the colour-major CSR of P A P^T and the split L / U index
// arrays of the factor.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "order.hpp"

namespace vfvm {

namespace {

// spreads the low 21 bits of x to every third bit
uint64_t spread3(uint64_t x) {
    x &= 0x1fffffULL;
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

}  // namespace

uint64_t morton3(const double* p, const double* lo, const double* hi) {
    uint64_t q[3];
    for (int a = 0; a < 3; ++a) {
        const double ext = hi[a] - lo[a];
        double f = ext > 0.0 ? (p[a] - lo[a]) / ext : 0.0;
        f = std::min(std::max(f, 0.0), 1.0);
        q[a] = (uint64_t)std::llround(f * 2097151.0);
    }
    return spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
}

int ilu_order_build(int64_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& cl, const double* pos,
                    int64_t nvert, IluOrder& out, std::string& err) {
    if (n < 0 || (int64_t)rp.size() != n + 1 || (int64_t)cl.size() != (int64_t)rp[n]) {
        err = "ILU(0): malformed pattern";
        return 1;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (cl[k] < 0 || cl[k] >= n) { err = "ILU(0): column index out of range"; return 1; }
    // symmetrised adjacency (rows of one colour must not touch each other in A or A^T)
    std::vector<int32_t> tcnt(n + 1, 0);
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) tcnt[cl[k] + 1]++;
    for (int64_t i = 0; i < n; ++i) tcnt[i + 1] += tcnt[i];
    std::vector<int32_t> trows(tcnt[n]);
    {
        std::vector<int32_t> fill(tcnt.begin(), tcnt.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) trows[fill[cl[k]]++] = (int32_t)i;
    }
    // deduplicated symmetrised adjacency without the diagonal (columns of a row ascending)
    std::vector<int32_t> aptr(n + 1, 0), adj;
    adj.reserve((size_t)rp[n] * 2);
    for (int64_t i = 0; i < n; ++i) {
        int32_t a = rp[i], ae = rp[i + 1], b = tcnt[i], be = tcnt[i + 1];
        int32_t last = -1;
        while (a < ae || b < be) {
            int32_t j;
            if (b >= be || (a < ae && cl[a] <= trows[b])) j = cl[a++];
            else j = trows[b++];
            if (j != (int32_t)i && j != last) { adj.push_back(j); last = j; }
        }
        aptr[i + 1] = (int32_t)adj.size();
    }
    // colouring order: smallest-last, which needs fewer colours than the natural order on
    // Delaunay graphs (10 against 11 on the 100^3 lattice)
    std::vector<int32_t> order(n);
    {
        std::vector<int32_t> deg(n);
        int32_t maxd = 0;
        for (int64_t i = 0; i < n; ++i) { deg[i] = aptr[i + 1] - aptr[i]; maxd = std::max(maxd, deg[i]); }
        std::vector<std::vector<int32_t>> bucket(maxd + 1);
        for (int64_t i = n - 1; i >= 0; --i) bucket[deg[i]].push_back((int32_t)i);
        std::vector<uint8_t> gone(n, 0);
        int32_t cur = 0;
        for (int64_t cnt = n - 1; cnt >= 0; --cnt) {
            int32_t v = -1;
            for (;;) {
                while (bucket[cur].empty()) ++cur;
                v = bucket[cur].back();
                bucket[cur].pop_back();
                if (!gone[v] && deg[v] == cur) break;  // skip stale entries
            }
            gone[v] = 1;
            order[cnt] = v;
            for (int32_t k = aptr[v]; k < aptr[v + 1]; ++k) {
                const int32_t u = adj[k];
                if (!gone[u]) bucket[--deg[u]].push_back(u);
            }
            cur = std::max(0, cur - 1);
        }
    }
    // greedy colouring in that order
    std::vector<int32_t> color(n, -1);
    std::vector<int64_t> mark(64, -1);
    int32_t ncol = 0;
    for (int64_t oi = 0; oi < n; ++oi) {
        const int32_t i = order[oi];
        for (int32_t k = aptr[i]; k < aptr[i + 1]; ++k) {
            const int32_t cj = color[adj[k]];
            if (cj >= 0) {
                if (cj >= (int32_t)mark.size()) mark.resize(cj + 1, -1);
                mark[cj] = oi;
            }
        }
        int32_t cc = 0;
        while (cc < (int32_t)mark.size() && mark[cc] == oi) ++cc;
        color[i] = cc;
        ncol = std::max(ncol, cc + 1);
    }
    // spatial key of every row: the Morton code of its vertex (rows of a coupled system are
    // blocks of nvert rows), or the row index without coordinates
    std::vector<uint64_t> key(n);
    if (pos && nvert > 0 && n % nvert == 0) {
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t v = 0; v < nvert; ++v)
            for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], pos[3 * v + a]); hi[a] = std::max(hi[a], pos[3 * v + a]); }
        for (int64_t i = 0; i < n; ++i) key[i] = morton3(pos + 3 * (i % nvert), lo, hi);
    } else {
        for (int64_t i = 0; i < n; ++i) key[i] = (uint64_t)i;
    }
    // permutation: rows sorted by (colour, key, index)
    std::vector<int32_t>& perm = out.perm;
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
        if (color[a] != color[b]) return color[a] < color[b];
        if (key[a] != key[b]) return key[a] < key[b];
        return a < b;
    });
    std::vector<int32_t> iperm(n);
    for (int64_t t = 0; t < n; ++t) iperm[perm[t]] = (int32_t)t;
    out.color_ptr.assign(ncol + 1, 0);
    for (int64_t i = 0; i < n; ++i) out.color_ptr[color[i] + 1]++;
    for (int32_t q = 0; q < ncol; ++q) out.color_ptr[q + 1] += out.color_ptr[q];
    // colour-major CSR, columns permuted and ascending; the split triangles
    const int64_t nnz = rp[n];
    out.prp.assign(n + 1, 0);
    out.pcol.resize(nnz);
    out.src.resize(nnz);
    out.pdiag.assign(n, -1);
    out.lptr.assign(n + 1, 0);
    out.uptr.assign(n + 1, 0);
    out.lpos.clear(); out.lcol.clear(); out.upos.clear(); out.ucol.clear();
    std::vector<std::pair<int32_t, int32_t>> tmp;
    for (int64_t t = 0; t < n; ++t) {
        const int32_t i = perm[t];
        tmp.clear();
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) tmp.push_back({iperm[cl[k]], k});
        std::sort(tmp.begin(), tmp.end());
        const int32_t base = out.prp[t];
        for (size_t q = 0; q < tmp.size(); ++q) {
            const int32_t j = tmp[q].first, p = base + (int32_t)q;
            out.pcol[p] = j;
            out.src[p] = tmp[q].second;
            if (j == t) out.pdiag[t] = p;
            else if (j < t) { out.lpos.push_back(p); out.lcol.push_back(j); }
            else { out.upos.push_back(p); out.ucol.push_back(j); }
        }
        out.prp[t + 1] = base + (int32_t)tmp.size();
        out.lptr[t + 1] = (int32_t)out.lpos.size();
        out.uptr[t + 1] = (int32_t)out.upos.size();
        if (out.pdiag[t] < 0) { err = "ILU(0): a row has no diagonal entry"; return 1; }
    }
    return 0;
}

}  // namespace vfvm
