// Jittered unit-cube lattice meshes.
//
// The reference places boundary nodes on a face grid and interior points by
// Poisson-disk sampling (src/base/nds.rs:48-88, src/base/pts.rs:190-385); the
// lattice meshes here are a different vertex set with the same data model:
// points first, then nodes, each node on exactly one boundary, and every vertex
// of the convex hull a node (src/base/geom_tet.rs:118-120).
//
// Vertex (i,j,k), 0 <= i,j,k < n, sits at (i,j,k)*h, h = 1/(n-1), plus a jitter
// of jitter*h*u per free axis, u = 2*U - 1 with U the top 53 bits of
//   splitmix64(splitmix64(splitmix64(splitmix64(splitmix64(seed) ^ i) ^ j) ^ k) ^ axis) / 2^53.
// Free axes: all three inside, the two in-plane axes on a face, the edge axis on
// a cube edge, none at a corner.  Coordinates are then rounded to multiples of
// 2^-22 so that the Delaunay predicates are exact.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "status.hpp"

namespace vfvm {

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

double lattice_jitter_unit(uint64_t seed, int64_t i, int64_t j, int64_t k, int axis) {
    uint64_t s = splitmix64(seed);
    s = splitmix64(s ^ (uint64_t)i);
    s = splitmix64(s ^ (uint64_t)j);
    s = splitmix64(s ^ (uint64_t)k);
    s = splitmix64(s ^ (uint64_t)axis);
    double u = (double)(s >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

int lattice_mesh(int n, uint64_t seed, double jitter, Mesh& m, std::string& err) {
    const int32_t lo[3] = {0, 0, 0}, hi[3] = {n - 1, n - 1, n - 1};
    return lattice_box_mesh(n, seed, jitter, lo, hi, m, err);
}

// The vertices of the n^3 lattice with lo <= (i, j, k) <= hi (a box of the unit cube).  A
// coordinate on a face of the box is the unjittered lattice value (0 and 1 exactly on the
// cube's faces), so each face is a plane and two boxes sharing a face place their common
// vertices identically (the same hash of the global lattice index jitters the free axes).
int lattice_box_mesh(int n, uint64_t seed, double jitter, const int32_t lo[3], const int32_t hi[3], Mesh& m,
                     std::string& err) {
    if (n < 3 || n > 1024) { err = "lattice: n must be in [3, 1024]"; return VFVM_ERR_INVALID_ARGUMENT; }
    if (!(jitter >= 0.0 && jitter < 0.5)) { err = "lattice: jitter must be in [0, 0.5)"; return VFVM_ERR_INVALID_ARGUMENT; }
    for (int ax = 0; ax < 3; ++ax)
        if (lo[ax] < 0 || hi[ax] > n - 1 || hi[ax] - lo[ax] < 2) {
            err = "lattice box: need 0 <= lo < hi <= n - 1 with at least 3 vertices per axis";
            return VFVM_ERR_INVALID_ARGUMENT;
        }
    const int64_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    const int64_t ntot = nx * ny * nz;
    const double h = 1.0 / (double)(n - 1);
    const double Q = 4194304.0;  // 2^22
    const int64_t QI = int64_t(1) << 22;
    const int64_t ext[3] = {nx, ny, nz};
    // quantised coordinate of lattice plane idx (0 and 2^22 exactly on the cube's faces)
    auto plane = [&](int64_t idx) -> int64_t {
        if (idx == 0) return 0;
        if (idx == n - 1) return QI;
        return (int64_t)std::llround((double)idx / (double)(n - 1) * Q);
    };

    // box-relative index r = idx - lo; on the box's boundary when r is 0 or ext - 1
    auto on_b = [&](int64_t r, int ax) { return r == 0 || r == ext[ax] - 1; };
    std::vector<int32_t> vid_of(ntot);
    int64_t num_pts = (nx - 2) * (ny - 2) * (nz - 2);
    int64_t p = 0, q = num_pts;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t li = i + nx * (j + ny * k);
                bool b = on_b(i, 0) || on_b(j, 1) || on_b(k, 2);
                vid_of[li] = (int32_t)(b ? q++ : p++);
            }
    m = Mesh();
    m.num_pts = num_pts;
    m.num_nds = ntot - num_pts;
    m.pos.assign(3 * ntot, 0.0);
    std::vector<int64_t> ixyz(3 * ntot);
    m.bnd_name = {"x0", "x1", "y0", "y1", "z0", "z1"};
    std::vector<std::vector<int32_t>> bl(6);
    m.nd_bnd.assign(m.num_nds, -1);
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t li = i + nx * (j + ny * k);
                int64_t v = vid_of[li];
                const int64_t rel[3] = {i, j, k};
                const int64_t idx[3] = {lo[0] + i, lo[1] + j, lo[2] + k};
                for (int ax = 0; ax < 3; ++ax) {
                    int64_t X;
                    if (on_b(rel[ax], ax)) {
                        X = plane(idx[ax]);
                    } else {
                        double x = (double)idx[ax] / (double)(n - 1);
                        x += jitter * h * lattice_jitter_unit(seed, idx[0], idx[1], idx[2], ax);
                        X = (int64_t)std::llround(x * Q);
                        const int64_t xlo = plane(lo[ax]) + 1, xhi = plane(hi[ax]) - 1;
                        if (X < xlo) X = xlo;
                        if (X > xhi) X = xhi;
                    }
                    ixyz[3 * v + ax] = X;
                    m.pos[3 * v + ax] = (double)X / Q;
                }
                if (v >= num_pts) {
                    int bid = on_b(i, 0) ? (i == 0 ? 0 : 1) : on_b(j, 1) ? (j == 0 ? 2 : 3) : (k == 0 ? 4 : 5);
                    int32_t nid = (int32_t)(v - num_pts);
                    bl[bid].push_back(nid);
                    m.nd_bnd[nid] = bid;
                }
            }
    m.bnd_ptr.assign(7, 0);
    for (int b = 0; b < 6; ++b) {
        m.bnd_ptr[b + 1] = m.bnd_ptr[b] + (int64_t)bl[b].size();
        m.bnd_nid.insert(m.bnd_nid.end(), bl[b].begin(), bl[b].end());
    }

    // insertion order: four corners spanning a volume, the other corners, then
    // coarse-to-fine lattice levels (stride 2^s), lattice order within a level
    std::vector<int32_t> order;
    order.reserve(ntot);
    std::vector<uint8_t> done(ntot, 0);
    auto push = [&](int64_t i, int64_t j, int64_t k) {
        int64_t li = i + nx * (j + ny * k);
        if (!done[li]) { done[li] = 1; order.push_back(vid_of[li]); }
    };
    const int64_t ex = nx - 1, ey = ny - 1, ez = nz - 1;
    push(0, 0, 0); push(ex, 0, 0); push(0, ey, 0); push(0, 0, ez);
    push(ex, ey, 0); push(ex, 0, ez); push(0, ey, ez); push(ex, ey, ez);
    const int64_t emax = std::max(ex, std::max(ey, ez));
    int64_t stride = 1;
    while (stride * 2 <= emax) stride *= 2;
    for (; stride >= 1; stride /= 2)
        for (int64_t k = 0; k < nz; k += stride)
            for (int64_t j = 0; j < ny; j += stride)
                for (int64_t i = 0; i < nx; i += stride) push(i, j, k);

    int st = delaunay_int(ixyz, order, m.tets, err);
    if (st != VFVM_OK) return st;
    return VFVM_OK;
}

}  // namespace vfvm
