// Exact Delaunay tetrahedralisation of points with integer coordinates.
//
// Bowyer-Watson insertion as in src/base/delaunay.rs:6-170 (visibility walk from
// the last new tetrahedron, cavity of circumspheres that strictly contain the new
// point, cavity replaced by the fan from the point), with two design changes:
//   * the convex hull is closed by ghost tetrahedra sharing one vertex at infinity
//     instead of a finite super-tetrahedron, so hull faces that hold coplanar
//     points (the faces of a cube) come out exactly;
//   * coordinates are integers below 2^22, so orient3d and insphere are evaluated
//     exactly in 128-bit integer arithmetic (|insphere| < 2^117) instead of by
//     adaptive floating-point expansions.
// A ghost (a,b,c,inf) conflicts with p if p lies strictly beyond its hull facet,
// or on the facet's plane strictly inside the facet's circumcircle; the latter is
// decided by the finite tetrahedron across the facet, whose circumsphere meets
// the plane in that circle.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "mesh.hpp"
#include "status.hpp"

namespace vfvm {

namespace {

typedef __int128 i128;
constexpr int INF = -1;
constexpr int64_t COORD_LIMIT = int64_t(1) << 22;

struct DT {
    const int64_t* X = nullptr;          // 3 per vertex
    std::vector<int32_t> tv;             // 4 per tet, vertex ids (INF only at slot 3)
    std::vector<int32_t> tn;             // 4 per tet, neighbour opposite slot i
    std::vector<uint8_t> alive;
    std::vector<int32_t> free_list;
    std::vector<uint32_t> mark;          // cavity stamp
    std::vector<uint32_t> seen;          // tested-not-in-conflict stamp
    uint32_t stamp = 0;
    uint64_t rng = 0x2545F4914F6CDD1Dull;
    int32_t last = 0;

    uint32_t rnd() { rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17; return (uint32_t)rng; }

    int32_t new_tet() {
        int32_t t;
        if (!free_list.empty()) { t = free_list.back(); free_list.pop_back(); }
        else {
            t = (int32_t)alive.size();
            tv.resize(tv.size() + 4); tn.resize(tn.size() + 4);
            alive.push_back(0); mark.push_back(0); seen.push_back(0);
        }
        alive[t] = 1;
        return t;
    }
    bool ghost(int32_t t) const { return tv[4 * t + 3] == INF; }

    // det[b-a; c-a; d-a] sign
    int orient(int a, int b, int c, int d) const {
        const int64_t *A = X + 3 * (int64_t)a, *B = X + 3 * (int64_t)b, *C = X + 3 * (int64_t)c, *D = X + 3 * (int64_t)d;
        int64_t bx = B[0] - A[0], by = B[1] - A[1], bz = B[2] - A[2];
        int64_t cx = C[0] - A[0], cy = C[1] - A[1], cz = C[2] - A[2];
        int64_t dx = D[0] - A[0], dy = D[1] - A[1], dz = D[2] - A[2];
        i128 det = (i128)bx * (cy * dz - cz * dy) - (i128)by * (cx * dz - cz * dx) + (i128)bz * (cx * dy - cy * dx);
        return (det > 0) - (det < 0);
    }
    static i128 det3(const int64_t* u, const int64_t* v, const int64_t* w) {
        return (i128)u[0] * (v[1] * w[2] - v[2] * w[1]) - (i128)u[1] * (v[0] * w[2] - v[2] * w[0]) +
               (i128)u[2] * (v[0] * w[1] - v[1] * w[0]);
    }
    // > 0 iff p strictly inside the circumsphere of positively oriented (a,b,c,d)
    int insphere(int a, int b, int c, int d, int p) const {
        const int64_t* P = X + 3 * (int64_t)p;
        int64_t q[4][3];
        int64_t l[4];
        int vs[4] = {a, b, c, d};
        for (int r = 0; r < 4; ++r) {
            const int64_t* V = X + 3 * (int64_t)vs[r];
            for (int k = 0; k < 3; ++k) q[r][k] = V[k] - P[k];
            l[r] = q[r][0] * q[r][0] + q[r][1] * q[r][1] + q[r][2] * q[r][2];
        }
        i128 D = -(i128)l[0] * det3(q[1], q[2], q[3]) + (i128)l[1] * det3(q[0], q[2], q[3]) -
                 (i128)l[2] * det3(q[0], q[1], q[3]) + (i128)l[3] * det3(q[0], q[1], q[2]);
        // D < 0 means inside for positive orientation
        return (D < 0) - (D > 0);
    }
    bool finite_conflict(int32_t t, int p) const {
        const int32_t* v = &tv[4 * t];
        return insphere(v[0], v[1], v[2], v[3], p) > 0;
    }
    bool conflict(int32_t t, int p) const {
        const int32_t* v = &tv[4 * t];
        if (v[3] != INF) return insphere(v[0], v[1], v[2], v[3], p) > 0;
        int o = orient(v[0], v[1], v[2], p);
        if (o > 0) return true;
        if (o < 0) return false;
        return finite_conflict(tn[4 * t + 3], p);
    }
    // orientation of tet t with vertex slot i replaced by p (t finite)
    int orient_replaced(int32_t t, int i, int p) const {
        int v[4] = {tv[4 * t], tv[4 * t + 1], tv[4 * t + 2], tv[4 * t + 3]};
        v[i] = p;
        return orient(v[0], v[1], v[2], v[3]);
    }

    int32_t locate(int p, std::string& err) {
        int32_t t = last;
        if (!alive[t]) {
            for (t = 0; t < (int32_t)alive.size() && !alive[t]; ++t) {}
        }
        int64_t steps = 0;
        const int64_t max_steps = 50000000;
        while (!ghost(t)) {
            uint32_t r = rnd() & 3;
            int32_t next = -1;
            for (int f = 0; f < 4; ++f) {
                int i = (int)((r + f) & 3);
                if (orient_replaced(t, i, p) < 0) { next = tn[4 * t + i]; break; }
            }
            if (next < 0) return t;
            t = next;
            if (++steps > max_steps) { err = "delaunay: point location did not terminate"; return -1; }
        }
        return t;  // ghost: p strictly outside the hull
    }

    struct BFace { int32_t t; int i; };

    int insert(int p, std::string& err) {
        int32_t t0 = locate(p, err);
        if (t0 < 0) return VFVM_ERR_TETRA_DEGENERATE_INPUT;
        if (!conflict(t0, p)) {
            err = "delaunay: located tetrahedron not in conflict (duplicate point?)";
            return VFVM_ERR_TETRA_DEGENERATE_INPUT;
        }
        std::vector<int32_t>& cav = cav_buf; cav.clear();
        std::vector<BFace>& bnd = bnd_buf;
        ++stamp;
        if (stamp == 0) { std::fill(mark.begin(), mark.end(), 0); std::fill(seen.begin(), seen.end(), 0); stamp = 1; }
        mark[t0] = stamp; cav.push_back(t0);
        for (;;) {
            bnd.clear();
            for (size_t k = 0; k < cav.size(); ++k) {
                int32_t t = cav[k];
                for (int i = 0; i < 4; ++i) {
                    int32_t nb = tn[4 * t + i];
                    if (mark[nb] == stamp) continue;
                    if (seen[nb] != stamp) {
                        if (conflict(nb, p)) { mark[nb] = stamp; cav.push_back(nb); continue; }
                        seen[nb] = stamp;
                    }
                    bnd.push_back({t, i});
                }
            }
            // safety net of delaunay.rs:15: every new tetrahedron must be proper
            bool grown = false;
            for (const BFace& bf : bnd) {
                int32_t t = bf.t;
                bool g = ghost(t);
                bool ok;
                if (!g) ok = orient_replaced(t, bf.i, p) > 0;
                else if (bf.i == 3) {
                    const int32_t* v = &tv[4 * t];
                    ok = orient(v[0], v[1], v[2], p) > 0;
                } else {
                    // new ghost facet: the two remaining finite vertices and p must not be collinear
                    int w[3]; int c = 0;
                    for (int k = 0; k < 3; ++k) w[c++] = (k == bf.i) ? p : tv[4 * t + k];
                    const int64_t *A = X + 3 * (int64_t)w[0], *B = X + 3 * (int64_t)w[1], *C = X + 3 * (int64_t)w[2];
                    int64_t ux = B[0] - A[0], uy = B[1] - A[1], uz = B[2] - A[2];
                    int64_t vx = C[0] - A[0], vy = C[1] - A[1], vz = C[2] - A[2];
                    ok = (uy * vz - uz * vy) != 0 || (uz * vx - ux * vz) != 0 || (ux * vy - uy * vx) != 0;
                }
                if (!ok) {
                    int32_t nb = tn[4 * t + bf.i];
                    if (mark[nb] != stamp) { mark[nb] = stamp; cav.push_back(nb); grown = true; }
                }
            }
            if (!grown) break;
            ++forced;
        }
        // create the fan
        std::vector<int32_t>& made = made_buf; made.clear();
        for (const BFace& bf : bnd) {
            int32_t t = bf.t;
            int32_t nt = new_tet();
            for (int k = 0; k < 4; ++k) { tv[4 * nt + k] = tv[4 * t + k]; tn[4 * nt + k] = -1; }
            tv[4 * nt + bf.i] = p;
            // a finite tet made from a ghost's facet: p replaced the infinite vertex
            int32_t outside = tn[4 * t + bf.i];
            tn[4 * nt + bf.i] = outside;
            for (int k = 0; k < 4; ++k)
                if (tn[4 * outside + k] == t) { tn[4 * outside + k] = nt; break; }
            made.push_back(nt);
        }
        // link new tets to each other across faces that contain p
        edge_map.clear();
        for (int32_t nt : made) {
            int ip = -1;
            for (int k = 0; k < 4; ++k) if (tv[4 * nt + k] == p) ip = k;
            for (int k = 0; k < 4; ++k) {
                if (k == ip) continue;
                // face opposite slot k contains p and the two vertices other than slots k, ip
                int a = -2, b = -2;
                for (int s = 0; s < 4; ++s) {
                    if (s == k || s == ip) continue;
                    if (a == -2) a = tv[4 * nt + s]; else b = tv[4 * nt + s];
                }
                if (a > b) std::swap(a, b);
                uint64_t key = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
                auto it = edge_map.find(key);
                if (it == edge_map.end()) edge_map.emplace(key, std::make_pair(nt, k));
                else {
                    tn[4 * nt + k] = it->second.first;
                    tn[4 * it->second.first + it->second.second] = nt;
                    edge_map.erase(it);
                }
            }
        }
        if (!edge_map.empty()) { err = "delaunay: cavity boundary is not a closed surface"; return VFVM_ERR_TETRA_DEGENERATE_INPUT; }
        for (int32_t t : cav) { alive[t] = 0; free_list.push_back(t); }
        last = made.back();
        if (ghost(last)) {
            for (int32_t nt : made) if (!ghost(nt)) { last = nt; break; }
        }
        return VFVM_OK;
    }

    std::vector<int32_t> cav_buf, made_buf;
    std::vector<BFace> bnd_buf;
    std::unordered_map<uint64_t, std::pair<int32_t, int>> edge_map;
    int64_t forced = 0;
};

}  // namespace

// ixyz: 3 per vertex, integers in [0, 2^22]; order: insertion order (all vertices,
// the first four must not be coplanar).  Output: finite tets in the reference's
// orientation (p3 below the counterclockwise face p0 p1 p2; geom_tet.rs:13).
int delaunay_int(const std::vector<int64_t>& ixyz, const std::vector<int32_t>& order,
                 std::vector<int32_t>& tets_out, std::string& err) {
    const int64_t nv = (int64_t)ixyz.size() / 3;
    if (nv < 4 || (int64_t)order.size() != nv) { err = "delaunay: need at least four vertices and a full insertion order"; return VFVM_ERR_TETRA_DEGENERATE_INPUT; }
    for (int64_t k = 0; k < 3 * nv; ++k)
        if (ixyz[k] < 0 || ixyz[k] > COORD_LIMIT) { err = "delaunay: integer coordinate outside [0, 2^22]"; return VFVM_ERR_INVALID_ARGUMENT; }
    DT dt;
    dt.X = ixyz.data();
    dt.tv.reserve(4 * 7 * nv); dt.tn.reserve(4 * 7 * nv);
    int a = order[0], b = order[1], c = order[2], d = order[3];
    int o = dt.orient(a, b, c, d);
    if (o == 0) { err = "delaunay: first four vertices are coplanar"; return VFVM_ERR_TETRA_DEGENERATE_INPUT; }
    if (o < 0) std::swap(a, b);
    int32_t t0 = dt.new_tet();
    int v0[4] = {a, b, c, d};
    for (int k = 0; k < 4; ++k) dt.tv[4 * t0 + k] = v0[k];
    // ghosts on the four faces: facet (f0,f1,f2) ordered so that the opposite finite vertex is on the negative side
    for (int i = 0; i < 4; ++i) {
        int f[3]; int cnt = 0;
        for (int k = 0; k < 4; ++k) if (k != i) f[cnt++] = v0[k];
        if (dt.orient(f[0], f[1], f[2], v0[i]) > 0) std::swap(f[1], f[2]);
        int32_t g = dt.new_tet();
        dt.tv[4 * g] = f[0]; dt.tv[4 * g + 1] = f[1]; dt.tv[4 * g + 2] = f[2]; dt.tv[4 * g + 3] = INF;
        dt.tn[4 * g + 3] = t0;
        dt.tn[4 * t0 + i] = g;
    }
    // ghost-ghost adjacency across faces (x, y, inf)
    {
        std::unordered_map<uint64_t, std::pair<int32_t, int>> m;
        for (int32_t g = 1; g <= 4; ++g)
            for (int k = 0; k < 3; ++k) {
                int x = -2, y = -2;
                for (int s = 0; s < 3; ++s) { if (s == k) continue; if (x == -2) x = dt.tv[4 * g + s]; else y = dt.tv[4 * g + s]; }
                if (x > y) std::swap(x, y);
                uint64_t key = ((uint64_t)(uint32_t)x << 32) | (uint32_t)y;
                auto it = m.find(key);
                if (it == m.end()) m.emplace(key, std::make_pair(g, k));
                else { dt.tn[4 * g + k] = it->second.first; dt.tn[4 * it->second.first + it->second.second] = g; }
            }
    }
    dt.last = t0;
    for (int64_t k = 4; k < nv; ++k) {
        int st = dt.insert(order[k], err);
        if (st != VFVM_OK) return st;
    }
    tets_out.clear();
    const int32_t nt = (int32_t)dt.alive.size();
    for (int32_t t = 0; t < nt; ++t) {
        if (!dt.alive[t] || dt.ghost(t)) continue;
        const int32_t* v = &dt.tv[4 * t];
        // internal orientation det[b-a;c-a;d-a] > 0; the reference wants the opposite sign
        tets_out.push_back(v[0]); tets_out.push_back(v[2]); tets_out.push_back(v[1]); tets_out.push_back(v[3]);
    }
    return VFVM_OK;
}

}  // namespace vfvm
