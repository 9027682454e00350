// Voronoi dual of a Delaunay tetrahedralisation built on the GPU: the same arrays as
// build_dual (dual.cpp, restating Tetrahedron::update_tet, src/base/geom_tet.rs:58-293),
// bitwise equal to them.
//
// Every accumulator of the reference is a sum in tetrahedron order (face pieces of an edge,
// control-volume centroid moments of a vertex) or in neighbour order (volumes, least-squares
// moments).  Here each sum is a gather: one thread walks the ascending list of the
// tetrahedra around its vertex (or edge) and adds the pieces in that order, so the result
// does not depend on scheduling; the expressions are those of dual.cpp operand by operand
// (compiled without contraction; division and square root are correctly rounded in
// double precision on CDNA, as on the host).  No sort of all faces is needed: a face is on
// the surface when no other tetrahedron around one of its vertices contains it.
//
// Data on the device: tets (4T), vertex -> tetrahedra lists (CSR, 4T entries, ascending),
// face pieces per (tetrahedron, edge, side) = (area, centroid) (48 doubles per tetrahedron),
// the neighbour CSR and the per-edge / per-vertex results, copied back into the host mesh.
// The boundary sub-faces (surface only) are built on the host (dual_subfaces).  The flow
// operators' pressure-gradient coefficients (grad_coeff, geom_tet.rs:295-347) are built the
// same way (build_grad_coeff_device), bitwise equal to grad_coeff.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "status.hpp"

namespace vfvm {

namespace {

constexpr int BLK = 256;
inline unsigned grid_of(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

struct D3 { double x, y, z; };
__device__ inline D3 add3(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 sub3(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 scale3(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 cross3(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline double len3(D3 a) { return sqrt(dot3(a, a)); }
__device__ inline D3 P(const double* __restrict__ pos, int64_t v) { return {pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]}; }

// src/base/delaunay.rs:204-211
__device__ inline D3 circumcircle3(D3 a, D3 b, D3 c) {
    D3 u = sub3(b, a), v = sub3(c, a);
    D3 w = cross3(u, v);
    D3 num = add3(scale3(dot3(u, u), cross3(v, w)), scale3(dot3(v, v), cross3(w, u)));
    return add3(a, scale3(0.5 / dot3(w, w), num));
}
// src/base/delaunay.rs:213-218
__device__ inline D3 circumsphere3(D3 a, D3 b, D3 c, D3 d) {
    D3 u = sub3(b, a), v = sub3(c, a), w = sub3(d, a);
    D3 num = add3(add3(scale3(dot3(u, u), cross3(v, w)), scale3(dot3(v, v), cross3(w, u))), scale3(dot3(w, w), cross3(u, v)));
    return add3(a, scale3(0.5 / dot3(u, cross3(v, w)), num));
}

// local vertex pairs of the six edges of a tetrahedron, in the reference's order
__constant__ int kEdges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__device__ inline int edge_index(int la, int lb) {  // la < lb
    return la == 0 ? lb - 1 : (la == 1 ? lb + 1 : 5);
}

// ---- vertex -> tetrahedra lists
__global__ void k_count(int64_t n4, const int32_t* __restrict__ tv, int32_t* __restrict__ cnt) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k < n4) atomicAdd(&cnt[tv[k]], 1);
}
__global__ void k_fill(int64_t n4, const int32_t* __restrict__ tv, const int32_t* __restrict__ ptr, int32_t* __restrict__ fill,
                       int32_t* __restrict__ lst) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k >= n4) return;
    const int32_t v = tv[k];
    const int32_t at = atomicAdd(&fill[v], 1);
    lst[ptr[v] + at] = (int32_t)(k / 4);
}
// each vertex's list ascending (insertion sort; lists hold a few tens of entries)
__global__ void k_sort_lists(int64_t n, const int32_t* __restrict__ ptr, int32_t* __restrict__ lst) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= n) return;
    const int32_t b = ptr[v], e = ptr[v + 1];
    for (int32_t i = b + 1; i < e; ++i) {
        const int32_t x = lst[i];
        int32_t j = i - 1;
        while (j >= b && lst[j] > x) { lst[j + 1] = lst[j]; --j; }
        lst[j + 1] = x;
    }
}
__device__ inline bool tet_has(const int32_t* __restrict__ q, int32_t v) { return q[0] == v || q[1] == v || q[2] == v || q[3] == v; }

// ---- surface faces (geom_tet.rs:99-120): face k of t (opposite local vertex k) is on the
// surface when no other tetrahedron around its first vertex contains all three of its
// vertices; flag 1 = surface, 2 = shared by more than two tetrahedra
__global__ void k_surface(int64_t n4, const int32_t* __restrict__ tv, const int32_t* __restrict__ ptr,
                          const int32_t* __restrict__ lst, uint8_t* __restrict__ flag) {
    const int64_t fk = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (fk >= n4) return;
    const int64_t t = fk / 4;
    const int k = (int)(fk % 4);
    const int32_t a = tv[4 * t + (k + 1) % 4], b = tv[4 * t + (k + 2) % 4], c = tv[4 * t + (k + 3) % 4];
    int others = 0;
    for (int32_t i = ptr[a]; i < ptr[a + 1]; ++i) {
        const int32_t u = lst[i];
        if (u == (int32_t)t) continue;
        const int32_t* q = tv + 4 * (int64_t)u;
        if (tet_has(q, b) && tet_has(q, c)) ++others;
    }
    flag[fk] = others == 0 ? 1 : (others > 1 ? 2 : 0);
}

// ---- neighbour CSR (ascending per row): the other vertices of the tetrahedra around v,
// sorted and de-duplicated in the vertex's own scratch segment (3 per listed tetrahedron)
__global__ void k_ngh_count(int64_t n, const int32_t* __restrict__ tv, const int32_t* __restrict__ ptr,
                            const int32_t* __restrict__ lst, int32_t* __restrict__ scr, int32_t* __restrict__ deg) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= n) return;
    const int32_t b = ptr[v], e = ptr[v + 1];
    int32_t* s = scr + 3 * (int64_t)b;
    int32_t m = 0;
    for (int32_t i = b; i < e; ++i) {
        const int32_t* q = tv + 4 * (int64_t)lst[i];
        for (int l = 0; l < 4; ++l) {
            const int32_t u = q[l];
            if (u == (int32_t)v) continue;
            // insert u into the sorted prefix s[0, m) unless present
            int32_t j = m - 1;
            while (j >= 0 && s[j] > u) --j;
            if (j >= 0 && s[j] == u) continue;
            for (int32_t r = m; r > j + 1; --r) s[r] = s[r - 1];
            s[j + 1] = u;
            ++m;
        }
    }
    deg[v] = m;
}
__global__ void k_ngh_fill(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ scr,
                           const int32_t* __restrict__ rp, int32_t* __restrict__ ngh) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= n) return;
    const int32_t* s = scr + 3 * (int64_t)ptr[v];
    for (int32_t i = rp[v]; i < rp[v + 1]; ++i) ngh[i] = s[i - rp[v]];
}

// ---- face pieces (geom_tet.rs:122-171): per tetrahedron, edge and side, the signed area
// of the Voronoi face piece and its centroid; stored (area, cx, cy, cz) at (12 t + 2 ed + side)
__global__ void k_pieces(int64_t T, const int32_t* __restrict__ tv, const double* __restrict__ pos, double* __restrict__ pc) {
    const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= T) return;
    const int32_t* q = tv + 4 * t;
    const D3 c_t = circumsphere3(P(pos, q[0]), P(pos, q[1]), P(pos, q[2]), P(pos, q[3]));
    for (int ed = 0; ed < 6; ++ed) {
        const int li = kEdges[ed][0], lj = kEdges[ed][1];
        int32_t ot[2];
        int c = 0;
        for (int l = 0; l < 4; ++l)
            if (l != li && l != lj) ot[c++] = q[l];
        const D3 pi = P(pos, q[li]), pj = P(pos, q[lj]);
        const double dist = len3(sub3(pj, pi));
        const D3 e = scale3(1.0 / dist, sub3(pj, pi));
        const D3 mm = scale3(0.5, add3(pi, pj));
        for (int side = 0; side < 2; ++side) {
            const D3 pa = P(pos, side == 0 ? ot[0] : ot[1]), pb = P(pos, side == 0 ? ot[1] : ot[0]);
            const D3 c_f = circumcircle3(pi, pj, pa);
            D3 u = sub3(sub3(pa, mm), scale3(dot3(sub3(pa, mm), e), e));
            u = scale3(1.0 / len3(u), u);
            D3 nv = cross3(e, u);
            if (dot3(sub3(pb, mm), nv) < 0.0) nv = scale3(-1.0, nv);
            const double s1 = dot3(sub3(c_f, mm), u);
            const double s2 = dot3(sub3(c_t, c_f), nv);
            const double ar = 0.5 * s1 * s2;
            const D3 cent = scale3(1.0 / 3.0, add3(mm, add3(c_f, c_t)));
            double* o = pc + 4 * (12 * t + 2 * ed + side);
            o[0] = ar; o[1] = cent.x; o[2] = cent.y; o[3] = cent.z;
        }
    }
}

// ---- per-edge sums of the pieces, in tetrahedron order: area and area-weighted skew of
// the undirected edge (a, b), a < b, stored at the slot of a -> b
__global__ void k_edge_sums(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ngh,
                            const int32_t* __restrict__ tv, const int32_t* __restrict__ ptr, const int32_t* __restrict__ lst,
                            const double* __restrict__ pos, const double* __restrict__ pc, int64_t E,
                            double* __restrict__ e_area, double* __restrict__ e_sk) {
    const int64_t a = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (a >= n) return;
    const D3 pa = P(pos, a);
    for (int32_t s = rp[a]; s < rp[a + 1]; ++s) {
        const int32_t b = ngh[s];
        if (b < a) continue;
        const D3 pb = P(pos, b);
        double ar_acc = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
        for (int32_t i = ptr[a]; i < ptr[a + 1]; ++i) {
            const int64_t t = lst[i];
            const int32_t* q = tv + 4 * t;
            int la = -1, lb = -1;
            for (int l = 0; l < 4; ++l) {
                if (q[l] == (int32_t)a) la = l;
                if (q[l] == b) lb = l;
            }
            if (lb < 0) continue;
            const int li = la < lb ? la : lb, lj = la < lb ? lb : la;
            const int ed = edge_index(li, lj);
            // midpoint of the tetrahedron's edge (q[li], q[lj]); the sum is symmetric
            const D3 mm = scale3(0.5, add3(li == la ? pa : pb, lj == la ? pa : pb));
            for (int side = 0; side < 2; ++side) {
                const double* o = pc + 4 * (12 * t + 2 * ed + side);
                const double ar = o[0];
                const D3 cent = {o[1], o[2], o[3]};
                ar_acc += ar;
                const D3 d = scale3(ar, sub3(cent, mm));
                sx += d.x; sy += d.y; sz += d.z;
            }
        }
        e_area[s] = ar_acc;
        e_sk[s] = sx; e_sk[E + s] = sy; e_sk[2 * E + s] = sz;
    }
}

// ---- control-volume centroid moments of vertex v (geom_tet.rs:160-168), in tetrahedron,
// edge, side order
__global__ void k_cent_sums(int64_t n, const int32_t* __restrict__ tv, const int32_t* __restrict__ ptr,
                            const int32_t* __restrict__ lst, const double* __restrict__ pos, const double* __restrict__ pc,
                            double* __restrict__ cent_sum) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= n) return;
    const D3 pv = P(pos, v);
    D3 acc = {0.0, 0.0, 0.0};
    for (int32_t i = ptr[v]; i < ptr[v + 1]; ++i) {
        const int64_t t = lst[i];
        const int32_t* q = tv + 4 * t;
        for (int ed = 0; ed < 6; ++ed) {
            const int li = kEdges[ed][0], lj = kEdges[ed][1];
            if (q[li] != (int32_t)v && q[lj] != (int32_t)v) continue;
            const D3 pi = P(pos, q[li]), pj = P(pos, q[lj]);
            const double dist = len3(sub3(pj, pi));
            for (int side = 0; side < 2; ++side) {
                const double* o = pc + 4 * (12 * t + 2 * ed + side);
                const double ar = o[0];
                const D3 cent = {o[1], o[2], o[3]};
                const double vol = ar * dist / 6.0;
                acc = add3(acc, scale3(vol * 0.75, sub3(cent, pv)));
            }
        }
    }
    cent_sum[v] = acc.x; cent_sum[n + v] = acc.y; cent_sum[2 * n + v] = acc.z;
}

__device__ inline int32_t slot_of(const int32_t* __restrict__ rp, const int32_t* __restrict__ ngh, int32_t a, int32_t b) {
    int32_t lo = rp[a], hi = rp[a + 1];
    while (lo < hi) {
        const int32_t mid = lo + (hi - lo) / 2;
        if (ngh[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---- per-directed-edge data (geom_tet.rs:173-201), written from the smaller vertex
__global__ void k_edge_data(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ngh,
                            const double* __restrict__ pos, const double* __restrict__ e_area, const double* __restrict__ e_sk,
                            int64_t E, double* __restrict__ dist_o, double* __restrict__ area_o, double* __restrict__ norm_o,
                            double* __restrict__ skew_o) {
    const int64_t a = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (a >= n) return;
    const D3 pa = P(pos, a);
    for (int32_t s = rp[a]; s < rp[a + 1]; ++s) {
        const int32_t b = ngh[s];
        if (b < a) continue;
        const int32_t r = slot_of(rp, ngh, b, (int32_t)a);
        const D3 pb = P(pos, b);
        const double dist = len3(sub3(pb, pa));
        const D3 nrm = scale3(1.0 / dist, sub3(pb, pa));
        const double ar = e_area[s];
        D3 sk = {0.0, 0.0, 0.0};
        if (fabs(ar) > 1e-12 * dist * dist) sk = scale3(1.0 / ar, D3{e_sk[s], e_sk[E + s], e_sk[2 * E + s]});
        const D3 nn = scale3(-1.0, nrm);
        for (int side = 0; side < 2; ++side) {
            const int64_t x = side == 0 ? s : r;
            const D3 nv = side == 0 ? nrm : nn;
            dist_o[x] = dist; area_o[x] = ar;
            norm_o[x] = nv.x; norm_o[E + x] = nv.y; norm_o[2 * E + x] = nv.z;
            skew_o[x] = sk.x; skew_o[E + x] = sk.y; skew_o[2 * E + x] = sk.z;
        }
    }
}

// ---- control volumes over the row in neighbour order and the control-volume skewness
// (geom_tet.rs:198-206)
__global__ void k_volumes(int64_t n, const int32_t* __restrict__ rp, const double* __restrict__ area, const double* __restrict__ dist,
                          const double* __restrict__ cent_sum, double* __restrict__ vol, double* __restrict__ cskew) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= n) return;
    double acc = 0.0;
    for (int32_t s = rp[v]; s < rp[v + 1]; ++s) acc += area[s] * dist[s] / 6.0;
    vol[v] = acc;
    const D3 c = scale3(1.0 / acc, D3{cent_sum[v], cent_sum[n + v], cent_sum[2 * n + v]});
    cskew[v] = c.x; cskew[n + v] = c.y; cskew[2 * n + v] = c.z;
}

// ---- least-squares gradient weights (geom_tet.rs:261-293); a singular system records the
// smallest such vertex in *bad
__global__ void k_lsq(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ngh, const double* __restrict__ pos,
                      int64_t E, double* __restrict__ gw, int32_t* __restrict__ bad) {
    const int64_t vc = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (vc >= n) return;
    const D3 pc = P(pos, vc);
    double mm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int32_t s = rp[vc]; s < rp[vc + 1]; ++s) {
        const D3 dd = sub3(P(pos, ngh[s]), pc);
        const double d[3] = {dd.x, dd.y, dd.z};
        const double omega = 1.0 / (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) mm[r][q] += omega * d[r] * d[q];
    }
    const double adj[3][3] = {
        {mm[1][1] * mm[2][2] - mm[1][2] * mm[2][1], mm[0][2] * mm[2][1] - mm[0][1] * mm[2][2], mm[0][1] * mm[1][2] - mm[0][2] * mm[1][1]},
        {mm[1][2] * mm[2][0] - mm[1][0] * mm[2][2], mm[0][0] * mm[2][2] - mm[0][2] * mm[2][0], mm[0][2] * mm[1][0] - mm[0][0] * mm[1][2]},
        {mm[1][0] * mm[2][1] - mm[1][1] * mm[2][0], mm[0][1] * mm[2][0] - mm[0][0] * mm[2][1], mm[0][0] * mm[1][1] - mm[0][1] * mm[1][0]},
    };
    const double det = mm[0][0] * adj[0][0] + mm[0][1] * adj[1][0] + mm[0][2] * adj[2][0];
    const double tr = mm[0][0] + mm[1][1] + mm[2][2];
    if (!(det > 1e-12 * (tr * tr * tr))) {
        atomicMin(bad, (int32_t)vc);
        return;
    }
    for (int32_t s = rp[vc]; s < rp[vc + 1]; ++s) {
        const D3 d = sub3(P(pos, ngh[s]), pc);
        const double omega = 1.0 / dot3(d, d);
        for (int r = 0; r < 3; ++r) gw[r * E + s] = omega * (adj[r][0] * d.x + adj[r][1] * d.y + adj[r][2] * d.z) / det;
    }
}


// ---- pressure-gradient coefficients (geom_tet.rs:295-347), restating grad_coeff.cpp as a
// gather.  The coefficients of vertex v are listed for every vertex of its 2-ring
// R2(v) = {v} u ngh(v) u ngh(ngh(v)), ascending.  Coefficient (v, m) is the sum, in the
// reference's push order, of the terms pushed for m: per face e = (v, n) the arithmetic
// means (to v, to n), then for s = v and s = n the skew-correction pairs (to ngh(s)[q], to s)
// in q order, then per boundary sub-face the three interpolation terms.  For m != s only
// the pair whose neighbour is m contributes (found by binary search in the ascending row of
// s); for m == s every pair's second term does.  The first term starts the sum (no 0 + x).
__global__ void k_r2_build(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ngh,
                           const int32_t* __restrict__ cap, int32_t* __restrict__ scr, int32_t* __restrict__ cnt) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= n) return;
    int32_t* s = scr + cap[v];
    int32_t m = 0;
    auto insert = [&](int32_t u) {
        int32_t j = m - 1;
        while (j >= 0 && s[j] > u) --j;
        if (j >= 0 && s[j] == u) return;
        for (int32_t r = m; r > j + 1; --r) s[r] = s[r - 1];
        s[j + 1] = u;
        ++m;
    };
    insert((int32_t)v);
    for (int32_t e = rp[v]; e < rp[v + 1]; ++e) {
        const int32_t nb = ngh[e];
        insert(nb);
        for (int32_t q = rp[nb]; q < rp[nb + 1]; ++q) insert(ngh[q]);
    }
    cnt[v] = m;
}
__global__ void k_r2_copy(int64_t n, const int32_t* __restrict__ cap, const int32_t* __restrict__ scr,
                          const int32_t* __restrict__ gp, int32_t* __restrict__ gvid, int32_t* __restrict__ grow) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= n) return;
    for (int32_t k = gp[v]; k < gp[v + 1]; ++k) {
        gvid[k] = scr[cap[v] + (k - gp[v])];
        grow[k] = (int32_t)v;
    }
}
struct GcAcc {
    double x = 0.0, y = 0.0, z = 0.0;
    bool first = true;
    __device__ void add(double a, double b, double c) {
        if (first) { x = a; y = b; z = c; first = false; }
        else { x += a; y += b; z += c; }
    }
};
__global__ void k_gc_coeff(int64_t G, const int32_t* __restrict__ grow, const int32_t* __restrict__ gvid,
                           const int32_t* __restrict__ rp, const int32_t* __restrict__ ngh, int64_t E,
                           const double* __restrict__ area_a, const double* __restrict__ norm, const double* __restrict__ skew,
                           const double* __restrict__ gw, const int32_t* __restrict__ bsp, const int32_t* __restrict__ bvj,
                           const int32_t* __restrict__ bvk, int64_t S, const double* __restrict__ barea,
                           const double* __restrict__ bnorm, const double* __restrict__ bwgt, double* __restrict__ gc) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k >= G) return;
    const int32_t v = grow[k], m = gvid[k];
    GcAcc acc;
    for (int32_t e = rp[v]; e < rp[v + 1]; ++e) {
        const int32_t nb = ngh[e];
        const double area = area_a[e];
        const double nx = norm[e], ny = norm[E + e], nz = norm[2 * E + e];
        const double sx = skew[e], sy = skew[E + e], sz = skew[2 * E + e];
        if (m == v) acc.add(0.5 * area * nx, 0.5 * area * ny, 0.5 * area * nz);
        if (m == nb) acc.add(0.5 * area * nx, 0.5 * area * ny, 0.5 * area * nz);
        for (int side = 0; side < 2; ++side) {
            const int32_t sv = side == 0 ? v : nb;
            if (m == sv) {
                for (int32_t q = rp[sv]; q < rp[sv + 1]; ++q) {
                    const double w_skew = 0.5 * (gw[q] * sx + gw[E + q] * sy + gw[2 * E + q] * sz) * area;
                    acc.add(-w_skew * nx, -w_skew * ny, -w_skew * nz);
                }
            } else {
                const int32_t q = slot_of(rp, ngh, sv, m);
                if (q < rp[sv + 1] && ngh[q] == m) {
                    const double w_skew = 0.5 * (gw[q] * sx + gw[E + q] * sy + gw[2 * E + q] * sz) * area;
                    acc.add(w_skew * nx, w_skew * ny, w_skew * nz);
                }
            }
        }
    }
    for (int32_t b = bsp[v]; b < bsp[v + 1]; ++b) {
        const double area = barea[b];
        const double nx = bnorm[b], ny = bnorm[S + b], nz = bnorm[2 * S + b];
        const double wv = bwgt[b], wj = bwgt[S + b], wk = bwgt[2 * S + b];
        if (m == v) acc.add(wv * area * nx, wv * area * ny, wv * area * nz);
        if (m == bvj[b]) acc.add(wj * area * nx, wj * area * ny, wj * area * nz);
        if (m == bvk[b]) acc.add(wk * area * nx, wk * area * ny, wk * area * nz);
    }
    gc[k] = acc.x; gc[G + k] = acc.y; gc[2 * G + k] = acc.z;
}

template <class T>
struct Buf {
    T* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)); }
};

#define DUAL_TRY(expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) { err = std::string("dual (device): ") + hipGetErrorString(e_); return VFVM_ERR_HIP; } \
    } while (0)

// exclusive scan of n counts into n + 1 offsets on the host; false when the total exceeds 2^31 - 1
bool scan_counts(const std::vector<int32_t>& cnt, std::vector<int32_t>& ptr) {
    ptr.assign(cnt.size() + 1, 0);
    int64_t acc = 0;
    for (size_t i = 0; i < cnt.size(); ++i) {
        acc += cnt[i];
        if (acc > INT32_MAX) return false;
        ptr[i + 1] = (int32_t)acc;
    }
    return true;
}

}  // namespace

int build_dual_device(Mesh& m, int device, std::string& err) {
    const int64_t N = m.num_tot();
    const int64_t T = (int64_t)m.tets.size() / 4;
    if (T == 0) { err = "dual: no tetrahedra"; return VFVM_ERR_TETRA_DEGENERATE_INPUT; }
    if (4 * T > INT32_MAX) { err = "dual: too many tetrahedra for 32-bit lists"; return VFVM_ERR_INVALID_ARGUMENT; }
    for (int64_t k = 0; k < 4 * T; ++k)
        if (m.tets[k] < 0 || m.tets[k] >= N) { err = "dual: tetrahedron vertex out of range"; return VFVM_ERR_INVALID_ARGUMENT; }
    int prev = 0;
    DUAL_TRY(hipGetDevice(&prev));
    DUAL_TRY(hipSetDevice(device));
    struct Restore { int d; ~Restore() { (void)hipSetDevice(d); } } restore{prev};
    hipStream_t st = nullptr;
    DUAL_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard { hipStream_t s; ~StreamGuard() { if (s) (void)hipStreamDestroy(s); } } sg{st};
    const int64_t n4 = 4 * T;

    Buf<int32_t> tv, cnt, ptr, fill, lst;
    Buf<double> pos;
    DUAL_TRY(tv.alloc(n4));
    DUAL_TRY(pos.alloc(3 * N));
    DUAL_TRY(cnt.alloc(N));
    DUAL_TRY(hipMemcpyAsync(tv.p, m.tets.data(), n4 * sizeof(int32_t), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemcpyAsync(pos.p, m.pos.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemsetAsync(cnt.p, 0, N * sizeof(int32_t), st));
    hipLaunchKernelGGL(k_count, dim3(grid_of(n4)), dim3(BLK), 0, st, n4, tv.p, cnt.p);
    std::vector<int32_t> h_cnt(N), h_ptr;
    DUAL_TRY(hipMemcpyAsync(h_cnt.data(), cnt.p, N * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipStreamSynchronize(st));
    scan_counts(h_cnt, h_ptr);  // total 4T < 2^31 (checked above)
    DUAL_TRY(ptr.alloc(N + 1));
    DUAL_TRY(fill.alloc(N));
    DUAL_TRY(lst.alloc(n4));
    DUAL_TRY(hipMemcpyAsync(ptr.p, h_ptr.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemsetAsync(fill.p, 0, N * sizeof(int32_t), st));
    hipLaunchKernelGGL(k_fill, dim3(grid_of(n4)), dim3(BLK), 0, st, n4, tv.p, ptr.p, fill.p, lst.p);
    hipLaunchKernelGGL(k_sort_lists, dim3(grid_of(N)), dim3(BLK), 0, st, N, ptr.p, lst.p);

    // surface faces -> host (geom_tet.rs:99-120), checks in the host build's order
    {
        Buf<uint8_t> flag;
        DUAL_TRY(flag.alloc(n4));
        hipLaunchKernelGGL(k_surface, dim3(grid_of(n4)), dim3(BLK), 0, st, n4, tv.p, ptr.p, lst.p, flag.p);
        std::vector<uint8_t> is_surf(n4);
        DUAL_TRY(hipMemcpyAsync(is_surf.data(), flag.p, n4, hipMemcpyDeviceToHost, st));
        DUAL_TRY(hipStreamSynchronize(st));
        for (int64_t k = 0; k < n4; ++k)
            if (is_surf[k] == 2) { err = "dual: a face is shared by more than two tetrahedra"; return VFVM_ERR_TETRA_DEGENERATE_INPUT; }
        std::vector<int32_t> surf = dual_surface_triangles(m, is_surf);
        if (int s = dual_check_surface(m, surf, err)) return s;
        dual_subfaces(m, surf);
    }

    // neighbour CSR
    Buf<int32_t> scr, deg, rp, ngh;
    DUAL_TRY(scr.alloc(3 * n4));
    DUAL_TRY(deg.alloc(N));
    hipLaunchKernelGGL(k_ngh_count, dim3(grid_of(N)), dim3(BLK), 0, st, N, tv.p, ptr.p, lst.p, scr.p, deg.p);
    std::vector<int32_t> h_deg(N);
    DUAL_TRY(hipMemcpyAsync(h_deg.data(), deg.p, N * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipStreamSynchronize(st));
    if (!scan_counts(h_deg, m.row_ptr)) { err = "dual: too many edges for 32-bit CSR"; return VFVM_ERR_INVALID_ARGUMENT; }
    const int64_t E = m.row_ptr[N];
    DUAL_TRY(rp.alloc(N + 1));
    DUAL_TRY(ngh.alloc(E));
    DUAL_TRY(hipMemcpyAsync(rp.p, m.row_ptr.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ngh_fill, dim3(grid_of(N)), dim3(BLK), 0, st, N, ptr.p, scr.p, rp.p, ngh.p);

    // face pieces and their sums
    Buf<double> pc, e_area, e_sk, cent;
    DUAL_TRY(pc.alloc(48 * T));
    DUAL_TRY(e_area.alloc(E));
    DUAL_TRY(e_sk.alloc(3 * E));
    DUAL_TRY(cent.alloc(3 * N));
    hipLaunchKernelGGL(k_pieces, dim3(grid_of(T)), dim3(BLK), 0, st, T, tv.p, pos.p, pc.p);
    hipLaunchKernelGGL(k_edge_sums, dim3(grid_of(N)), dim3(BLK), 0, st, N, rp.p, ngh.p, tv.p, ptr.p, lst.p, pos.p, pc.p, E,
                       e_area.p, e_sk.p);
    hipLaunchKernelGGL(k_cent_sums, dim3(grid_of(N)), dim3(BLK), 0, st, N, tv.p, ptr.p, lst.p, pos.p, pc.p, cent.p);

    // per-edge data, volumes, least squares
    Buf<double> dist, area, norm, skew, vol, cskew, gw;
    Buf<int32_t> bad;
    DUAL_TRY(dist.alloc(E));
    DUAL_TRY(area.alloc(E));
    DUAL_TRY(norm.alloc(3 * E));
    DUAL_TRY(skew.alloc(3 * E));
    DUAL_TRY(vol.alloc(N));
    DUAL_TRY(cskew.alloc(3 * N));
    DUAL_TRY(gw.alloc(3 * E));
    DUAL_TRY(bad.alloc(1));
    const int32_t no_bad = INT32_MAX;
    DUAL_TRY(hipMemcpyAsync(bad.p, &no_bad, sizeof(int32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_edge_data, dim3(grid_of(N)), dim3(BLK), 0, st, N, rp.p, ngh.p, pos.p, e_area.p, e_sk.p, E, dist.p, area.p,
                       norm.p, skew.p);
    hipLaunchKernelGGL(k_volumes, dim3(grid_of(N)), dim3(BLK), 0, st, N, rp.p, area.p, dist.p, cent.p, vol.p, cskew.p);
    hipLaunchKernelGGL(k_lsq, dim3(grid_of(N)), dim3(BLK), 0, st, N, rp.p, ngh.p, pos.p, E, gw.p, bad.p);
    DUAL_TRY(hipGetLastError());

    m.ngh.resize(E);
    m.dist.resize(E); m.area.resize(E); m.norm.resize(3 * E); m.skew.resize(3 * E); m.gw.resize(3 * E);
    m.vol.resize(N); m.cskew.resize(3 * N);
    int32_t h_bad = 0;
    DUAL_TRY(hipMemcpyAsync(&h_bad, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.ngh.data(), ngh.p, E * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.dist.data(), dist.p, E * sizeof(double), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.area.data(), area.p, E * sizeof(double), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.norm.data(), norm.p, 3 * E * sizeof(double), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.skew.data(), skew.p, 3 * E * sizeof(double), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.gw.data(), gw.p, 3 * E * sizeof(double), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.vol.data(), vol.p, N * sizeof(double), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.cskew.data(), cskew.p, 3 * N * sizeof(double), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipStreamSynchronize(st));
    if (h_bad != INT32_MAX) {
        err = "dual: singular least-squares gradient at vid " + std::to_string(h_bad);
        return VFVM_ERR_TETRA_SINGULAR_GRADIENT;
    }
    m.has_dual = true;
    return VFVM_OK;
}

int build_grad_coeff_device(Mesh& m, int device, std::string& err) {
    if (!m.has_dual) { err = "grad_coeff: mesh without dual"; return VFVM_ERR_INVALID_ARGUMENT; }
    if (m.has_grad_coeff) return VFVM_OK;
    const int64_t N = m.num_tot(), E = m.num_edges(), S = m.num_subfaces();
    // scratch capacity of each vertex's 2-ring: 1 + deg(v) + sum of its neighbours' degrees
    std::vector<int32_t> h_cap(N + 1, 0);
    {
        int64_t acc = 0;
        for (int64_t v = 0; v < N; ++v) {
            int64_t c = 1;
            for (int32_t e = m.row_ptr[v]; e < m.row_ptr[v + 1]; ++e) c += 1 + (m.row_ptr[m.ngh[e] + 1] - m.row_ptr[m.ngh[e]]);
            h_cap[v] = (int32_t)acc;
            acc += c;
            if (acc > INT32_MAX) { err = "grad_coeff: 2-ring scratch exceeds 2^31 entries"; return VFVM_ERR_INVALID_ARGUMENT; }
        }
        h_cap[N] = (int32_t)acc;
    }
    int prev = 0;
    DUAL_TRY(hipGetDevice(&prev));
    DUAL_TRY(hipSetDevice(device));
    struct Restore { int d; ~Restore() { (void)hipSetDevice(d); } } restore{prev};
    hipStream_t st = nullptr;
    DUAL_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard { hipStream_t s; ~StreamGuard() { if (s) (void)hipStreamDestroy(s); } } sg{st};

    Buf<int32_t> rp, ngh, cap, scr, cnt;
    DUAL_TRY(rp.alloc(N + 1));
    DUAL_TRY(ngh.alloc(E));
    DUAL_TRY(cap.alloc(N + 1));
    DUAL_TRY(scr.alloc(h_cap[N]));
    DUAL_TRY(cnt.alloc(N));
    DUAL_TRY(hipMemcpyAsync(rp.p, m.row_ptr.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemcpyAsync(ngh.p, m.ngh.data(), E * sizeof(int32_t), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemcpyAsync(cap.p, h_cap.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_r2_build, dim3(grid_of(N)), dim3(BLK), 0, st, N, rp.p, ngh.p, cap.p, scr.p, cnt.p);
    std::vector<int32_t> h_cnt(N);
    DUAL_TRY(hipMemcpyAsync(h_cnt.data(), cnt.p, N * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipStreamSynchronize(st));
    std::vector<int32_t> gptr;
    if (!scan_counts(h_cnt, gptr)) { err = "grad_coeff: more than 2^31 coefficients"; return VFVM_ERR_INVALID_ARGUMENT; }
    const int64_t G = gptr[N];
    Buf<int32_t> gp, gvid, grow, bsp, bvj, bvk;
    Buf<double> area, norm, skew, gw, barea, bnorm, bwgt, gc;
    DUAL_TRY(gp.alloc(N + 1));
    DUAL_TRY(gvid.alloc(G));
    DUAL_TRY(grow.alloc(G));
    DUAL_TRY(hipMemcpyAsync(gp.p, gptr.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_r2_copy, dim3(grid_of(N)), dim3(BLK), 0, st, N, cap.p, scr.p, gp.p, gvid.p, grow.p);
    DUAL_TRY(area.alloc(E));
    DUAL_TRY(norm.alloc(3 * E));
    DUAL_TRY(skew.alloc(3 * E));
    DUAL_TRY(gw.alloc(3 * E));
    DUAL_TRY(bsp.alloc(N + 1));
    DUAL_TRY(bvj.alloc(S));
    DUAL_TRY(bvk.alloc(S));
    DUAL_TRY(barea.alloc(S));
    DUAL_TRY(bnorm.alloc(3 * S));
    DUAL_TRY(bwgt.alloc(3 * S));
    DUAL_TRY(gc.alloc(3 * G));
    DUAL_TRY(hipMemcpyAsync(area.p, m.area.data(), E * sizeof(double), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemcpyAsync(norm.p, m.norm.data(), 3 * E * sizeof(double), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemcpyAsync(skew.p, m.skew.data(), 3 * E * sizeof(double), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemcpyAsync(gw.p, m.gw.data(), 3 * E * sizeof(double), hipMemcpyHostToDevice, st));
    DUAL_TRY(hipMemcpyAsync(bsp.p, m.bsf_ptr.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (S > 0) {
        DUAL_TRY(hipMemcpyAsync(bvj.p, m.bsf_vj.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
        DUAL_TRY(hipMemcpyAsync(bvk.p, m.bsf_vk.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
        DUAL_TRY(hipMemcpyAsync(barea.p, m.bsf_area.data(), S * sizeof(double), hipMemcpyHostToDevice, st));
        DUAL_TRY(hipMemcpyAsync(bnorm.p, m.bsf_norm.data(), 3 * S * sizeof(double), hipMemcpyHostToDevice, st));
        DUAL_TRY(hipMemcpyAsync(bwgt.p, m.bsf_wgt.data(), 3 * S * sizeof(double), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_gc_coeff, dim3(grid_of(G)), dim3(BLK), 0, st, G, grow.p, gvid.p, rp.p, ngh.p, E, area.p, norm.p, skew.p,
                       gw.p, bsp.p, bvj.p, bvk.p, S, barea.p, bnorm.p, bwgt.p, gc.p);
    DUAL_TRY(hipGetLastError());
    m.gc_ptr = std::move(gptr);
    m.gc_vid.resize(G);
    m.gc_c.resize(3 * G);
    DUAL_TRY(hipMemcpyAsync(m.gc_vid.data(), gvid.p, G * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipMemcpyAsync(m.gc_c.data(), gc.p, 3 * G * sizeof(double), hipMemcpyDeviceToHost, st));
    DUAL_TRY(hipStreamSynchronize(st));
    m.has_grad_coeff = true;
    return VFVM_OK;
}

}  // namespace vfvm
