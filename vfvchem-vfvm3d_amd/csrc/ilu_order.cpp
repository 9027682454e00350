// Multicolour ILU(0) ordering (host): greedy colouring of the symmetrised pattern graph in
// smallest-last order (Matula-Beck degeneracy order, reversed; Saad, Iterative Methods for
// Sparse Linear Systems, 2nd ed., section 12.4 multicolour orderings), rows sorted by (colour,
// Morton code of their vertex), the colour-major CSR of P A P^T and the split L / U index
// arrays of the factor.
#include <algorithm>
#include <cmath>
#include <cstdint>
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

// The tiled order (see IluOrder): pencils along the longest axis of the box, ILU_TILE_W mean
// spacings across, cut along the axis into segments of about ILU_TILE_TARGET rows; tiles
// coloured greedily in index order over the pattern's tile graph; inside a tile, levels by a
// greedy colouring of the tile's own graph in order of decreasing degree inside the tile.
// Fails (false) when a limit is exceeded; the caller then takes the multicolour order.
bool tiled_order(int64_t n, const std::vector<int32_t>& aptr, const std::vector<int32_t>& adj, const double* pos,
                 const double* lo, const double* hi, const std::vector<uint64_t>& key, IluOrder& out) {
    int ax = 0;
    for (int a = 1; a < 3; ++a)
        if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
    const int bx = (ax + 1) % 3, cx = (ax + 2) % 3;
    double vol = 1.0;
    for (int a = 0; a < 3; ++a) vol *= std::max(hi[a] - lo[a], 1e-300);
    const double hs = std::cbrt(vol / (double)n);  // mean spacing
    // per axis: about m = extent / hs vertices, spacing extent / (m - 1); a vertex's index on
    // the axis is its rounded coordinate in spacings (on a lattice: its lattice index), and a
    // tile spans ILU_TILE_W consecutive indices across the pencil
    int64_t m[3];
    double sp[3];
    for (int a = 0; a < 3; ++a) {
        m[a] = std::max<int64_t>(1, (int64_t)std::llround((hi[a] - lo[a]) / hs));
        sp[a] = m[a] > 1 ? (hi[a] - lo[a]) / (double)(m[a] - 1) : 1.0;
    }
    auto index = [&](double x, int a) {
        const int64_t q = (int64_t)std::llround((x - lo[a]) / sp[a]);
        return std::min(std::max<int64_t>(q, 0), m[a] - 1);
    };
    const int64_t w = (int64_t)ILU_TILE_W;
    const int64_t nb = (m[bx] + w - 1) / w, nc = (m[cx] + w - 1) / w;
    const double per_pencil = (double)n / (double)(nb * nc);
    const int64_t ns = std::max<int64_t>(1, (int64_t)std::ceil(per_pencil / ILU_TILE_TARGET));
    const int64_t seg_len = (m[ax] + ns - 1) / ns;
    const int64_t ntl = ns * nb * nc;
    if (ntl > INT32_MAX) return false;
    std::vector<int32_t> tile(n);
    for (int64_t i = 0; i < n; ++i) {
        const double* p = pos + 3 * i;
        tile[i] = (int32_t)(((index(p[ax], ax) / seg_len) * nb + index(p[bx], bx) / w) * nc + index(p[cx], cx) / w);
    }
    // tile graph (edges between different tiles), greedy colouring in tile index order
    std::vector<std::vector<int32_t>> tnb(ntl);
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = aptr[i]; k < aptr[i + 1]; ++k)
            if (tile[adj[k]] != tile[i]) tnb[tile[i]].push_back(tile[adj[k]]);
    std::vector<int32_t> tcol(ntl, -1), tcnt(ntl, 0);
    for (int64_t i = 0; i < n; ++i) tcnt[tile[i]]++;
    int32_t ntc = 0;
    for (int64_t t = 0; t < ntl; ++t) {
        if (tcnt[t] == 0) continue;
        if (tcnt[t] > ILU_TILE_RMAX) return false;
        uint32_t used = 0;
        for (int32_t u : tnb[t])
            if (tcol[u] >= 0 && tcol[u] < 32) used |= 1u << tcol[u];
        int32_t q = 0;
        while (q < 32 && (used >> q & 1u)) ++q;
        if (q >= ILU_MAXTC) return false;
        tcol[t] = q;
        ntc = std::max(ntc, q + 1);
    }
    // levels inside each tile: greedy colouring of the tile's graph, rows by decreasing degree
    // inside the tile (then Morton code)
    std::vector<std::vector<int32_t>> rows_of(ntl);
    for (int64_t i = 0; i < n; ++i) rows_of[tile[i]].push_back((int32_t)i);
    std::vector<int32_t> lev(n, -1), ideg(n, 0);
    int32_t nlev = 0;
    for (int64_t t = 0; t < ntl; ++t) {
        std::vector<int32_t>& rs = rows_of[t];
        for (int32_t i : rs)
            for (int32_t k = aptr[i]; k < aptr[i + 1]; ++k) ideg[i] += tile[adj[k]] == t;
        std::vector<int32_t> ord = rs;
        std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
            if (ideg[a] != ideg[b]) return ideg[a] > ideg[b];
            if (key[a] != key[b]) return key[a] < key[b];
            return a < b;
        });
        for (int32_t i : ord) {
            uint32_t used = 0;
            for (int32_t k = aptr[i]; k < aptr[i + 1]; ++k) {
                const int32_t j = adj[k];
                if (tile[j] == t && lev[j] >= 0 && lev[j] < 32) used |= 1u << lev[j];
            }
            int32_t l = 0;
            while (l < 32 && (used >> l & 1u)) ++l;
            if (l >= ILU_MAXLEV) return false;
            lev[i] = l;
            nlev = std::max(nlev, l + 1);
        }
    }
    // rows by (tile colour, tile, level, key)
    std::vector<int32_t>& perm = out.perm;
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
        if (tcol[tile[a]] != tcol[tile[b]]) return tcol[tile[a]] < tcol[tile[b]];
        if (tile[a] != tile[b]) return tile[a] < tile[b];
        if (lev[a] != lev[b]) return lev[a] < lev[b];
        if (key[a] != key[b]) return key[a] < key[b];
        return a < b;
    });
    out.nlev = nlev;
    out.tile_row.assign(1, 0);
    out.tc_ptr.assign(1, 0);
    out.color_ptr.assign(1, 0);
    out.lev_row.clear();
    out.rmax = 0;
    for (int64_t t = 0; t < n;) {
        const int32_t tl = tile[perm[t]];
        int64_t e = t;
        while (e < n && tile[perm[e]] == tl) ++e;
        const int64_t k = (int64_t)out.tile_row.size() - 1;
        // a new tile colour starts
        if (k > 0 && tcol[tile[perm[out.tile_row[k] - 1]]] != tcol[tl]) {
            out.tc_ptr.push_back((int32_t)k);
            out.color_ptr.push_back((int32_t)t);
        }
        // first row of each level (the tile's rows are sorted by level)
        int64_t r = t;
        for (int l = 0; l <= ILU_MAXLEV; ++l) {
            while (r < e && lev[perm[r]] < l) ++r;
            out.lev_row.push_back((int32_t)r);
        }
        out.tile_row.push_back((int32_t)e);
        out.rmax = std::max<int32_t>(out.rmax, (int32_t)(e - t));
        t = e;
    }
    out.tc_ptr.push_back((int32_t)out.tile_row.size() - 1);
    out.color_ptr.push_back((int32_t)n);
    // factorisation groups: every (tile colour, level)
    out.fac_ptr.assign(1, 0);
    for (size_t q = 0; q + 1 < out.tc_ptr.size(); ++q)
        for (int l = 0; l < nlev; ++l) {
            for (int32_t k = out.tc_ptr[q]; k < out.tc_ptr[q + 1]; ++k)
                for (int32_t r = out.lev_row[(size_t)k * (ILU_MAXLEV + 1) + l]; r < out.lev_row[(size_t)k * (ILU_MAXLEV + 1) + l + 1]; ++r)
                    out.fac_rows.push_back(r);
            out.fac_ptr.push_back((int32_t)out.fac_rows.size());
        }
    return true;
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
                    int64_t nvert, IluOrder& out, std::string& err, bool tiled) {
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
    // spatial key of every row: the Morton code of its vertex (rows of a coupled system are
    // blocks of nvert rows), or the row index without coordinates
    std::vector<uint64_t> key(n);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const bool has_pos = pos && nvert > 0 && n % nvert == 0;
    if (has_pos) {
        for (int64_t v = 0; v < nvert; ++v)
            for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], pos[3 * v + a]); hi[a] = std::max(hi[a], pos[3 * v + a]); }
        for (int64_t i = 0; i < n; ++i) key[i] = morton3(pos + 3 * (i % nvert), lo, hi);
    } else {
        for (int64_t i = 0; i < n; ++i) key[i] = (uint64_t)i;
    }
    out.tile_row.clear(); out.tc_ptr.clear(); out.lev_row.clear(); out.fac_rows.clear(); out.fac_ptr.clear();
    out.nlev = out.rmax = out.emax_l = out.emax_u = 0;
    if (!(tiled && has_pos && n == nvert && n > 0 && tiled_order(n, aptr, adj, pos, lo, hi, key, out))) {
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
    // permutation: rows sorted by (colour, key, index)
    std::vector<int32_t>& perm = out.perm;
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
        if (color[a] != color[b]) return color[a] < color[b];
        if (key[a] != key[b]) return key[a] < key[b];
        return a < b;
    });
    out.color_ptr.assign(ncol + 1, 0);
    for (int64_t i = 0; i < n; ++i) out.color_ptr[color[i] + 1]++;
    for (int32_t q = 0; q < ncol; ++q) out.color_ptr[q + 1] += out.color_ptr[q];
    }  // multicolour order
    std::vector<int32_t>& perm = out.perm;
    std::vector<int32_t> iperm(n);
    for (int64_t t = 0; t < n; ++t) iperm[perm[t]] = (int32_t)t;
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
    for (size_t k = 0; k + 1 < out.tile_row.size(); ++k) {
        const int32_t r0 = out.tile_row[k], r1 = out.tile_row[k + 1];
        out.emax_l = std::max(out.emax_l, out.lptr[r1] - out.lptr[r0]);
        out.emax_u = std::max(out.emax_u, out.uptr[r1] - out.uptr[r0]);
    }
    return 0;
}

bool tile_roles_build(const IluOrder& o, int max_deps, std::vector<int32_t>& role, std::vector<int32_t>& dep) {
    role.clear();
    dep.clear();
    const int64_t nt = (int64_t)o.tile_row.size() - 1;
    const int ntc = (int)o.tc_ptr.size() - 1;
    if (nt <= 0 || ntc <= 0) return false;
    const int64_t n = o.tile_row[nt];
    std::vector<int32_t> tile_of(n), colour(nt);
    for (int64_t k = 0; k < nt; ++k)
        for (int32_t r = o.tile_row[k]; r < o.tile_row[k + 1]; ++r) tile_of[r] = (int32_t)k;
    for (int q = 0; q < ntc; ++q)
        for (int32_t k = o.tc_ptr[q]; k < o.tc_ptr[q + 1]; ++k) colour[k] = q;
    // tile adjacency of the symmetrised pattern
    std::vector<std::vector<int32_t>> adj(nt);
    for (int64_t r = 0; r < n; ++r)
        for (int32_t e = o.prp[r]; e < o.prp[r + 1]; ++e) {
            const int32_t a = tile_of[r], b = tile_of[o.pcol[e]];
            if (a != b) {
                adj[a].push_back(b);
                adj[b].push_back(a);
            }
        }
    for (auto& v : adj) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    auto add = [&](int32_t k, int kind) -> bool {
        const size_t d0 = dep.size();
        const int q = colour[k];
        if (kind == 2) dep.push_back((k << 1) | 1);
        for (int32_t j : adj[k]) {
            if (colour[j] == q) return false;  // adjacent tiles of one colour: not a tile colouring
            if (kind != 2 && colour[j] < q) dep.push_back((j << 1) | 1);
            if (kind == 2 && colour[j] > q) dep.push_back(j << 1);
        }
        if ((int64_t)(dep.size() - d0) > max_deps) return false;
        const int32_t r0 = o.tile_row[k], r1 = o.tile_row[k + 1];
        const int32_t rec[TILE_ROLE_INTS] = {r0, r1, o.lptr[r0], o.lptr[r1], o.uptr[r0], o.uptr[r1], k, kind,
                                             (int32_t)d0, (int32_t)dep.size(), 0, 0, 0, 0, 0, 0};
        role.insert(role.end(), rec, rec + TILE_ROLE_INTS);
        return true;
    };
    for (int q = 0; q < ntc; ++q)
        for (int32_t k = o.tc_ptr[q]; k < o.tc_ptr[q + 1]; ++k)
            if (!add(k, q == ntc - 1 ? 3 : 1)) return false;
    for (int q = ntc - 2; q >= 0; --q)
        for (int32_t k = o.tc_ptr[q]; k < o.tc_ptr[q + 1]; ++k)
            if (!add(k, 2)) return false;
    return true;
}

}  // namespace vfvm
