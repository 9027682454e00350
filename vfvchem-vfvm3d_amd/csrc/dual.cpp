// Voronoi dual of a Delaunay tetrahedralisation, restating
// Tetrahedron::update_tet (src/base/geom_tet.rs:58-293) into flat CSR arrays.
//
// Arithmetic follows the reference expression by expression (the same operand
// order, no fused multiply-add) and every accumulator is summed in the
// reference's order: face pieces per edge in tetrahedron order (geom_tet.rs:131-171),
// control-volume volumes over a vertex's edges in ascending key order
// (geom_tet.rs:175-201), sub-faces in surface-triangle order (geom_tet.rs:213-259).
// The layout differs: directed edges are CSR rows with neighbours ascending
// (the order geom_tet.rs:174-196 produces), per-edge data are SoA arrays, and
// boundary sub-faces are a second CSR.
//
// The kept-tetrahedra filter (geom_tet.rs:90-97) keeps every tetrahedron here:
// the lattice domains are convex, so every Delaunay tetrahedron lies inside.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "status.hpp"

namespace vfvm {

namespace {

struct V3 { double x, y, z; };
inline V3 add3(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scale3(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross3(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len3(V3 a) { return std::sqrt(dot3(a, a)); }

// src/base/delaunay.rs:204-211
inline V3 circumcircle3(V3 a, V3 b, V3 c) {
    V3 u = sub3(b, a), v = sub3(c, a);
    V3 w = cross3(u, v);
    V3 num = add3(scale3(dot3(u, u), cross3(v, w)), scale3(dot3(v, v), cross3(w, u)));
    return add3(a, scale3(0.5 / dot3(w, w), num));
}
// src/base/delaunay.rs:213-218
inline V3 circumsphere3(V3 a, V3 b, V3 c, V3 d) {
    V3 u = sub3(b, a), v = sub3(c, a), w = sub3(d, a);
    V3 num = add3(add3(scale3(dot3(u, u), cross3(v, w)), scale3(dot3(v, v), cross3(w, u))), scale3(dot3(w, w), cross3(u, v)));
    return add3(a, scale3(0.5 / dot3(u, cross3(v, w)), num));
}

}  // namespace

// Surface triangles (faces of one tetrahedron, flagged per (t, k)) in tetrahedron order,
// counterclockwise seen from outside (geom_tet.rs:99-120): 3 vids per triangle.
std::vector<int32_t> dual_surface_triangles(const Mesh& m, const std::vector<uint8_t>& is_surf) {
    static const int SURF[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
    const int64_t T = (int64_t)m.tets.size() / 4;
    const int32_t* tv = m.tets.data();
    std::vector<int32_t> surf;
    for (int64_t t = 0; t < T; ++t)
        for (int k = 0; k < 4; ++k)
            if (is_surf[4 * t + k]) {
                surf.push_back(tv[4 * t + SURF[k][0]]);
                surf.push_back(tv[4 * t + SURF[k][1]]);
                surf.push_back(tv[4 * t + SURF[k][2]]);
            }
    return surf;
}

// Points (interior vertices) must not lie on the surface (geom_tet.rs:118-120).
int dual_check_surface(const Mesh& m, const std::vector<int32_t>& surf, std::string& err) {
    const int64_t NP = m.num_pts;
    std::vector<uint8_t> on(NP, 0);
    int64_t num_bnd = 0;
    for (int32_t v : surf) if (v < NP && !on[v]) { on[v] = 1; ++num_bnd; }
    if (num_bnd > 0) { err = "dual: " + std::to_string(num_bnd) + " points lie on the domain boundary"; return VFVM_ERR_TETRA_POINTS_ON_BOUNDARY; }
    return VFVM_OK;
}

// Boundary sub-faces (geom_tet.rs:208-259), in surface-triangle order.
void dual_subfaces(Mesh& m, const std::vector<int32_t>& surf) {
    const int64_t N = m.num_tot();
    auto P = [&](int64_t v) { return V3{m.pos[3 * v], m.pos[3 * v + 1], m.pos[3 * v + 2]}; };
    const int64_t NS = (int64_t)surf.size() / 3;
    m.bsf_ptr.assign(N + 1, 0);
    for (int64_t k = 0; k < 3 * NS; ++k) m.bsf_ptr[surf[k] + 1] += 1;
    for (int64_t v = 0; v < N; ++v) m.bsf_ptr[v + 1] += m.bsf_ptr[v];
    const int64_t S = m.bsf_ptr[N];
    m.bsf_vj.assign(S, 0); m.bsf_vk.assign(S, 0); m.bsf_area.assign(S, 0.0);
    m.bsf_norm.assign(3 * S, 0.0); m.bsf_wgt.assign(3 * S, 0.0);
    m.bnd_sum.assign(N, 0.0);
    {
        std::vector<int32_t> fill(m.bsf_ptr.begin(), m.bsf_ptr.end() - 1);
        for (int64_t f = 0; f < NS; ++f) {
            const int32_t* t = &surf[3 * f];
            V3 p0 = P(t[0]), p1 = P(t[1]), p2 = P(t[2]);
            V3 nrm = cross3(sub3(p1, p0), sub3(p2, p0));
            nrm = scale3(1.0 / len3(nrm), nrm);
            V3 c = circumcircle3(p0, p1, p2);
            double sq[3] = {dot3(sub3(p2, p1), sub3(p2, p1)), dot3(sub3(p0, p2), sub3(p0, p2)), dot3(sub3(p1, p0), sub3(p1, p0))};
            double lam_raw[3] = {sq[0] * (sq[1] + sq[2] - sq[0]), sq[1] * (sq[2] + sq[0] - sq[1]), sq[2] * (sq[0] + sq[1] - sq[2])};
            double lam_sum = lam_raw[0] + lam_raw[1] + lam_raw[2];
            double lam[3] = {lam_raw[0] / lam_sum, lam_raw[1] / lam_sum, lam_raw[2] / lam_sum};
            for (int loc = 0; loc < 3; ++loc) {
                int32_t vi = t[loc], vj = t[(loc + 1) % 3], vk = t[(loc + 2) % 3];
                V3 pi = P(vi), pj = P(vj), pk = P(vk);
                double ar[2];
                for (int side = 0; side < 2; ++side) {
                    V3 pa = side == 0 ? pj : pk, pb = side == 0 ? pk : pj;
                    V3 mm = scale3(0.5, add3(pi, pa));
                    V3 e = sub3(pa, pi);
                    double len = len3(e);
                    V3 u = sub3(sub3(pb, mm), scale3(dot3(sub3(pb, mm), e) / (len * len), e));
                    ar[side] = 0.25 * len * dot3(sub3(c, mm), scale3(1.0 / len3(u), u));
                }
                double area_sum = ar[0] + ar[1];
                double l0 = lam[loc], l1 = lam[(loc + 1) % 3], l2 = lam[(loc + 2) % 3];
                V3 cent_j = {(1.5 + l0) / 3.0, (0.5 + l1) / 3.0, l2 / 3.0};
                V3 cent_k = {(1.5 + l0) / 3.0, l1 / 3.0, (0.5 + l2) / 3.0};
                V3 w = {1.0, 0.0, 0.0};
                if (std::fabs(area_sum) > 1e-12 * len3(sub3(pj, pi)) * len3(sub3(pk, pi)))
                    w = scale3(1.0 / area_sum, add3(scale3(ar[0], cent_j), scale3(ar[1], cent_k)));
                m.bnd_sum[vi] += area_sum;
                int32_t x = fill[vi]++;
                m.bsf_vj[x] = vj; m.bsf_vk[x] = vk; m.bsf_area[x] = area_sum;
                m.bsf_norm[x] = nrm.x; m.bsf_norm[S + x] = nrm.y; m.bsf_norm[2 * S + x] = nrm.z;
                m.bsf_wgt[x] = w.x; m.bsf_wgt[S + x] = w.y; m.bsf_wgt[2 * S + x] = w.z;
            }
        }
    }

}

int build_dual(Mesh& m, int /*nthreads*/, std::string& err) {
    const int64_t N = m.num_tot();
    const int64_t T = (int64_t)m.tets.size() / 4;
    if (T == 0) { err = "dual: no tetrahedra"; return VFVM_ERR_TETRA_DEGENERATE_INPUT; }
    const int32_t* tv = m.tets.data();
    auto P = [&](int64_t v) { return V3{m.pos[3 * v], m.pos[3 * v + 1], m.pos[3 * v + 2]}; };
    for (int64_t k = 0; k < 4 * T; ++k)
        if (tv[k] < 0 || tv[k] >= N) { err = "dual: tetrahedron vertex out of range"; return VFVM_ERR_INVALID_ARGUMENT; }

    // ---- step 1: boundary faces = faces of exactly one tetrahedron (geom_tet.rs:99-120)
    struct FaceRec { int32_t a, b, c; int64_t tk; };
    std::vector<FaceRec> faces(4 * T);
    for (int64_t t = 0; t < T; ++t)
        for (int k = 0; k < 4; ++k) {
            int32_t f[3] = {tv[4 * t + (k + 1) % 4], tv[4 * t + (k + 2) % 4], tv[4 * t + (k + 3) % 4]};
            std::sort(f, f + 3);
            faces[4 * t + k] = {f[0], f[1], f[2], 4 * t + k};
        }
    std::sort(faces.begin(), faces.end(), [](const FaceRec& x, const FaceRec& y) {
        if (x.a != y.a) return x.a < y.a;
        if (x.b != y.b) return x.b < y.b;
        if (x.c != y.c) return x.c < y.c;
        return x.tk < y.tk;
    });
    std::vector<uint8_t> is_surf(4 * T, 0);
    for (size_t s = 0; s < faces.size();) {
        size_t e = s + 1;
        while (e < faces.size() && faces[e].a == faces[s].a && faces[e].b == faces[s].b && faces[e].c == faces[s].c) ++e;
        if (e - s == 1) is_surf[faces[s].tk] = 1;
        else if (e - s > 2) { err = "dual: a face is shared by more than two tetrahedra"; return VFVM_ERR_TETRA_DEGENERATE_INPUT; }
        s = e;
    }
    std::vector<FaceRec>().swap(faces);
    std::vector<int32_t> surf = dual_surface_triangles(m, is_surf);
    std::vector<uint8_t>().swap(is_surf);
    if (int st = dual_check_surface(m, surf, err)) return st;

    // ---- neighbour CSR, ascending per row
    std::vector<int64_t> cnt(N + 1, 0);
    for (int64_t k = 0; k < 4 * T; ++k) cnt[tv[k] + 1] += 3;
    for (int64_t v = 0; v < N; ++v) cnt[v + 1] += cnt[v];
    std::vector<int32_t> inc(cnt[N]);
    {
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (int64_t t = 0; t < T; ++t)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    if (a != b) inc[fill[tv[4 * t + a]]++] = tv[4 * t + b];
    }
    m.row_ptr.assign(N + 1, 0);
    std::vector<int32_t> ngh;
    ngh.reserve(cnt[N] / 4);
    for (int64_t v = 0; v < N; ++v) {
        auto b = inc.begin() + cnt[v], e = inc.begin() + cnt[v + 1];
        std::sort(b, e);
        auto u = std::unique(b, e);
        ngh.insert(ngh.end(), b, u);
        m.row_ptr[v + 1] = (int32_t)ngh.size();
    }
    std::vector<int32_t>().swap(inc);
    std::vector<int64_t>().swap(cnt);
    m.ngh = std::move(ngh);
    const int64_t E = (int64_t)m.ngh.size();
    if (E > INT32_MAX) { err = "dual: too many edges for 32-bit CSR"; return VFVM_ERR_INVALID_ARGUMENT; }
    auto slot = [&](int32_t a, int32_t b) -> int64_t {  // directed edge a -> b
        auto beg = m.ngh.begin() + m.row_ptr[a], end = m.ngh.begin() + m.row_ptr[a + 1];
        return std::lower_bound(beg, end, b) - m.ngh.begin();
    };

    // ---- step 2: Voronoi face pieces (geom_tet.rs:122-171)
    std::vector<double> e_area(E, 0.0), e_sk(3 * E, 0.0);  // indexed by the slot of (min -> max)
    std::vector<V3> cent_sum(N, V3{0.0, 0.0, 0.0});
    static const int EDGES[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int64_t t = 0; t < T; ++t) {
        const int32_t* q = tv + 4 * t;
        V3 c_t = circumsphere3(P(q[0]), P(q[1]), P(q[2]), P(q[3]));
        for (int ed = 0; ed < 6; ++ed) {
            int li = EDGES[ed][0], lj = EDGES[ed][1];
            int32_t vi = q[li], vj = q[lj];
            int32_t ot[2]; int c = 0;
            for (int l = 0; l < 4; ++l) if (l != li && l != lj) ot[c++] = q[l];
            V3 pi = P(vi), pj = P(vj);
            double dist = len3(sub3(pj, pi));
            V3 e = scale3(1.0 / dist, sub3(pj, pi));
            V3 mm = scale3(0.5, add3(pi, pj));
            int64_t s = slot(std::min(vi, vj), std::max(vi, vj));
            for (int side = 0; side < 2; ++side) {
                int32_t va = side == 0 ? ot[0] : ot[1];
                int32_t vb = side == 0 ? ot[1] : ot[0];
                V3 pa = P(va), pb = P(vb);
                V3 c_f = circumcircle3(pi, pj, pa);
                V3 u = sub3(sub3(pa, mm), scale3(dot3(sub3(pa, mm), e), e));
                u = scale3(1.0 / len3(u), u);
                V3 n = cross3(e, u);
                if (dot3(sub3(pb, mm), n) < 0.0) n = scale3(-1.0, n);
                double s1 = dot3(sub3(c_f, mm), u);
                double s2 = dot3(sub3(c_t, c_f), n);
                double ar = 0.5 * s1 * s2;
                V3 cent = scale3(1.0 / 3.0, add3(mm, add3(c_f, c_t)));
                e_area[s] += ar;
                V3 d = scale3(ar, sub3(cent, mm));
                e_sk[s] += d.x; e_sk[E + s] += d.y; e_sk[2 * E + s] += d.z;
                double vol = ar * dist / 6.0;
                cent_sum[vi] = add3(cent_sum[vi], scale3(vol * 0.75, sub3(cent, pi)));
                cent_sum[vj] = add3(cent_sum[vj], scale3(vol * 0.75, sub3(cent, pj)));
            }
        }
    }

    // ---- step 3: per-directed-edge data and CV volumes (geom_tet.rs:173-201)
    m.dist.assign(E, 0.0); m.area.assign(E, 0.0);
    m.norm.assign(3 * E, 0.0); m.skew.assign(3 * E, 0.0);
    m.vol.assign(N, 0.0);
    for (int64_t a = 0; a < N; ++a) {
        for (int64_t s = m.row_ptr[a]; s < m.row_ptr[a + 1]; ++s) {
            int32_t b = m.ngh[s];
            if (b < a) continue;  // handled from the smaller vertex
            int64_t r = slot(b, (int32_t)a);
            V3 pa = P(a), pb = P(b);
            double dist = len3(sub3(pb, pa));
            V3 nrm = scale3(1.0 / dist, sub3(pb, pa));
            double ar = e_area[s];
            V3 sk = V3{0.0, 0.0, 0.0};
            if (std::fabs(ar) > 1e-12 * dist * dist) sk = scale3(1.0 / ar, V3{e_sk[s], e_sk[E + s], e_sk[2 * E + s]});
            V3 nn = scale3(-1.0, nrm);
            for (int side = 0; side < 2; ++side) {
                int64_t x = side == 0 ? s : r;
                V3 nv = side == 0 ? nrm : nn;
                m.dist[x] = dist; m.area[x] = ar;
                m.norm[x] = nv.x; m.norm[E + x] = nv.y; m.norm[2 * E + x] = nv.z;
                m.skew[x] = sk.x; m.skew[E + x] = sk.y; m.skew[2 * E + x] = sk.z;
            }
        }
    }
    // volumes summed per vertex over its edges in ascending key order
    for (int64_t v = 0; v < N; ++v) {
        double acc = 0.0;
        for (int64_t s = m.row_ptr[v]; s < m.row_ptr[v + 1]; ++s) acc += m.area[s] * m.dist[s] / 6.0;
        m.vol[v] = acc;
    }
    std::vector<double>().swap(e_area);
    std::vector<double>().swap(e_sk);
    // control-volume skewness (geom_tet.rs:203-206)
    m.cskew.assign(3 * N, 0.0);
    for (int64_t v = 0; v < N; ++v) {
        V3 c = scale3(1.0 / m.vol[v], cent_sum[v]);
        m.cskew[v] = c.x; m.cskew[N + v] = c.y; m.cskew[2 * N + v] = c.z;
    }
    std::vector<V3>().swap(cent_sum);

    dual_subfaces(m, surf);
    std::vector<int32_t>().swap(surf);

    // ---- step 5: least-squares gradient weights (geom_tet.rs:261-293)
    m.gw.assign(3 * E, 0.0);
    for (int64_t vc = 0; vc < N; ++vc) {
        V3 pc = P(vc);
        double mm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int64_t s = m.row_ptr[vc]; s < m.row_ptr[vc + 1]; ++s) {
            V3 dd = sub3(P(m.ngh[s]), pc);
            double d[3] = {dd.x, dd.y, dd.z};
            double omega = 1.0 / (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) mm[r][q] += omega * d[r] * d[q];
        }
        double adj[3][3] = {
            {mm[1][1] * mm[2][2] - mm[1][2] * mm[2][1], mm[0][2] * mm[2][1] - mm[0][1] * mm[2][2], mm[0][1] * mm[1][2] - mm[0][2] * mm[1][1]},
            {mm[1][2] * mm[2][0] - mm[1][0] * mm[2][2], mm[0][0] * mm[2][2] - mm[0][2] * mm[2][0], mm[0][2] * mm[1][0] - mm[0][0] * mm[1][2]},
            {mm[1][0] * mm[2][1] - mm[1][1] * mm[2][0], mm[0][1] * mm[2][0] - mm[0][0] * mm[2][1], mm[0][0] * mm[1][1] - mm[0][1] * mm[1][0]},
        };
        double det = mm[0][0] * adj[0][0] + mm[0][1] * adj[1][0] + mm[0][2] * adj[2][0];
        double tr = mm[0][0] + mm[1][1] + mm[2][2];
        if (!(det > 1e-12 * (tr * tr * tr))) {
            err = "dual: singular least-squares gradient at vid " + std::to_string(vc);
            return VFVM_ERR_TETRA_SINGULAR_GRADIENT;
        }
        for (int64_t s = m.row_ptr[vc]; s < m.row_ptr[vc + 1]; ++s) {
            V3 d = sub3(P(m.ngh[s]), pc);
            double omega = 1.0 / dot3(d, d);
            for (int r = 0; r < 3; ++r)
                m.gw[r * E + s] = omega * (adj[r][0] * d.x + adj[r][1] * d.y + adj[r][2] * d.z) / det;
        }
    }
    m.has_dual = true;
    return VFVM_OK;
}

}  // namespace vfvm
