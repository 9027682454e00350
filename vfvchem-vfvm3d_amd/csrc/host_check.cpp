// Host-only exercise of the library's C++ host code — lattice generation, exact Delaunay
// insertion, the Voronoi dual, grad_coeff, the tiled and multicolour ILU(0) orderings and the CSV / VTU
// writers — for a build with AddressSanitizer and UndefinedBehaviorSanitizer (make sanitize).
// It needs no GPU: the device builders are replaced by functions that report their absence.
// Exits non-zero (or the sanitizers abort) on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "common.hpp"
#include "mesh.hpp"
#include "order.hpp"

namespace {
thread_local std::string g_err;
int fails = 0;
void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "host check FAILED: %s (%s)\n", what, g_err.c_str());
        ++fails;
    }
}
}  // namespace

int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int clear_error() {
    g_err.clear();
    return VFVM_OK;
}

namespace vfvm {
int build_dual_device(Mesh&, int, std::string& err) {
    err = "no GPU in the host check";
    return VFVM_ERR_HIP;
}
int build_grad_coeff_device(Mesh&, int, std::string& err) {
    err = "no GPU in the host check";
    return VFVM_ERR_HIP;
}
}  // namespace vfvm

// the scalar operator pattern {v} u ngh(v), ascending, of a mesh
static void scalar_pattern(const vfvm_mesh_info& info, const std::vector<int32_t>& rp, const std::vector<int32_t>& ngh,
                           std::vector<int32_t>& mrow, std::vector<int32_t>& mcol) {
    const int64_t N = info.num_pts + info.num_nds;
    mrow.assign(N + 1, 0);
    mcol.clear();
    for (int64_t v = 0; v < N; ++v) {
        bool placed = false;
        for (int32_t e = rp[v]; e < rp[v + 1]; ++e) {
            if (!placed && ngh[e] > v) { mcol.push_back((int32_t)v); placed = true; }
            mcol.push_back(ngh[e]);
        }
        if (!placed) mcol.push_back((int32_t)v);
        mrow[v + 1] = (int32_t)mcol.size();
    }
}

// The one-launch sweeps' roles: every tile swept forward once and backward once; every
// dependency an earlier role (so every wait ends); a role depends on every tile whose z it
// reads; a tile's backward sweep waits for the backward sweep of every tile whose forward
// sweep read its forward z; a tile doing both sweeps holds all its U entries.
static void check_tile_roles(const vfvm::IluOrder& o) {
    std::vector<int32_t> role, dep;
    check(vfvm::tile_roles_build(o, 512, role, dep), "tile_roles_build");
    const int64_t nt = (int64_t)o.tile_row.size() - 1, n = o.tile_row[nt];
    const int64_t nrole = (int64_t)role.size() / vfvm::TILE_ROLE_INTS;
    std::vector<int64_t> fwd(nt, -1), bwd(nt, -1);
    std::vector<int32_t> tile_of(n);
    for (int64_t k = 0; k < nt; ++k)
        for (int32_t r = o.tile_row[k]; r < o.tile_row[k + 1]; ++r) tile_of[r] = (int32_t)k;
    bool once = true, earlier = true, reads = true, war = true, inside = true;
    for (int64_t i = 0; i < nrole; ++i) {
        const int32_t* R = role.data() + i * vfvm::TILE_ROLE_INTS;
        const int32_t k = R[6];
        if (R[7] & 1) { once = once && fwd[k] < 0; fwd[k] = i; }
        if (R[7] & 2) { once = once && bwd[k] < 0; bwd[k] = i; }
        once = once && R[0] == o.tile_row[k] && R[1] == o.tile_row[k + 1];
    }
    for (int64_t k = 0; k < nt; ++k) once = once && fwd[k] >= 0 && bwd[k] >= 0;
    check(once, "every tile is swept forward once and backward once");
    if (!once) return;
    auto has = [&](const int32_t* R, int32_t d) {
        for (int32_t q = R[8]; q < R[9]; ++q)
            if (dep[q] == d) return true;
        return false;
    };
    for (int64_t i = 0; i < nrole; ++i) {
        const int32_t* R = role.data() + i * vfvm::TILE_ROLE_INTS;
        const int32_t k = R[6];
        for (int32_t q = R[8]; q < R[9]; ++q) {
            const int32_t j = dep[q] >> 1;
            earlier = earlier && ((dep[q] & 1) ? fwd[j] : bwd[j]) < i;
        }
        for (int32_t r = R[0]; r < R[1]; ++r) {
            if (R[7] & 1)
                for (int32_t e = o.lptr[r]; e < o.lptr[r + 1]; ++e) {
                    const int32_t j = tile_of[o.lcol[e]];
                    if (j == k) continue;
                    reads = reads && has(R, (j << 1) | 1);
                    // tile j's forward z is read here: j's backward sweep waits for this tile's
                    const int32_t* B = role.data() + bwd[j] * vfvm::TILE_ROLE_INTS;
                    war = war && has(B, k << 1);
                }
            if (R[7] == 3)
                for (int32_t e = o.uptr[r]; e < o.uptr[r + 1]; ++e) inside = inside && tile_of[o.ucol[e]] == k;
            if (R[7] == 2)
                for (int32_t e = o.uptr[r]; e < o.uptr[r + 1]; ++e) {
                    const int32_t j = tile_of[o.ucol[e]];
                    if (j != k) reads = reads && has(R, j << 1);
                }
        }
        if (R[7] == 2) reads = reads && has(R, (k << 1) | 1);
    }
    check(earlier, "every dependency of a role is an earlier role");
    check(reads, "a role waits for every tile whose z it reads");
    check(war, "a backward sweep waits for the forward sweeps that read its tile's z");
    check(inside, "a tile doing both sweeps holds all its U entries");
}

static void check_ilu_order(const std::vector<int32_t>& mrow, const std::vector<int32_t>& mcol, const std::vector<double>& pos,
                            int64_t nvert) {
    const int64_t n = (int64_t)mrow.size() - 1;
    for (bool tiled : {true, false}) {
        vfvm::IluOrder o;
        std::string err;
        check(vfvm::ilu_order_build(n, mrow, mcol, pos.data(), nvert, o, err, tiled) == 0, "ilu_order_build");
        std::vector<int> seen(n, 0), colour(n, -1), tile(n, -1), level(n, 0);
        for (int64_t t = 0; t < n; ++t) seen[o.perm[t]]++;
        bool bij = true;
        for (int64_t i = 0; i < n; ++i) bij = bij && seen[i] == 1;
        check(bij, "ILU permutation is a bijection");
        for (size_t q = 0; q + 1 < o.color_ptr.size(); ++q)
            for (int32_t t = o.color_ptr[q]; t < o.color_ptr[q + 1]; ++t) colour[o.perm[t]] = (int)q;
        // tiled order: a tile's rows and levels (a multicolour row is a tile of its own)
        for (int64_t t = 0; t < n; ++t) tile[o.perm[t]] = (int)t;
        for (size_t k = 0; k + 1 < o.tile_row.size(); ++k)
            for (int l = 0; l < vfvm::ILU_MAXLEV; ++l)
                for (int32_t t = o.lev_row[k * (vfvm::ILU_MAXLEV + 1) + l]; t < o.lev_row[k * (vfvm::ILU_MAXLEV + 1) + l + 1]; ++t) {
                    tile[o.perm[t]] = (int)k;
                    level[o.perm[t]] = l;
                }
        check(!tiled || nvert != n || !o.tile_row.empty(), "the tiled order is taken for a scalar system");
        bool indep = true;
        for (int64_t i = 0; i < n; ++i)
            for (int32_t k = mrow[i]; k < mrow[i + 1]; ++k) {
                const int32_t j = mcol[k];
                if (j != i && colour[j] == colour[i] && (tile[j] != tile[i] || level[j] == level[i])) indep = false;
            }
        check(indep, "rows of one colour are coupled only inside a tile, at different levels");
        check((int64_t)o.lpos.size() + (int64_t)o.upos.size() + n == (int64_t)mcol.size(), "split triangles cover the pattern");
        bool lower = true;
        for (int64_t t = 0; t < n; ++t)
            for (int32_t k = o.lptr[t]; k < o.lptr[t + 1]; ++k) lower = lower && o.lcol[k] < t;
        check(lower, "L entries lie below the diagonal");
        if (!o.tile_row.empty()) check_tile_roles(o);
    }
}

int main() {
    for (int n : {3, 6, 9, 22}) {
        for (uint64_t seed : {0ull, 5ull}) {
            vfvm_mesh* m = nullptr;
            check(vfvm_mesh_lattice(n, seed, 0.3, &m) == VFVM_OK, "vfvm_mesh_lattice");
            if (!m) continue;
            vfvm_mesh_info info;
            check(vfvm_mesh_info_get(m, &info) == VFVM_OK, "vfvm_mesh_info_get");
            const int64_t N = info.num_pts + info.num_nds;
            check(N == (int64_t)n * n * n, "vertex count");
            check(vfvm_mesh_build_grad_coeff(m) == VFVM_OK, "vfvm_mesh_build_grad_coeff");
            check(vfvm_mesh_info_get(m, &info) == VFVM_OK && info.num_grad_coeff > 0, "grad_coeff size");
            std::vector<int32_t> rp(N + 1), ngh(info.num_edges), tets(4 * info.num_tets), bnid(info.num_nds);
            std::vector<int64_t> bptr(info.num_bnd + 1);
            std::vector<double> pos(3 * N), vol(N), area(info.num_edges);
            check(vfvm_mesh_get(m, VFVM_MESH_ROW_PTR, rp.data()) == VFVM_OK, "get row_ptr");
            check(vfvm_mesh_get(m, VFVM_MESH_NGH, ngh.data()) == VFVM_OK, "get ngh");
            check(vfvm_mesh_get(m, VFVM_MESH_POS, pos.data()) == VFVM_OK, "get pos");
            check(vfvm_mesh_get(m, VFVM_MESH_VOL, vol.data()) == VFVM_OK, "get vol");
            check(vfvm_mesh_get(m, VFVM_MESH_TETS, tets.data()) == VFVM_OK, "get tets");
            check(vfvm_mesh_get(m, VFVM_MESH_BND_PTR, bptr.data()) == VFVM_OK, "get bnd_ptr");
            check(vfvm_mesh_get(m, VFVM_MESH_BND_NID, bnid.data()) == VFVM_OK, "get bnd_nid");
            double vsum = 0.0;
            for (double v : vol) vsum += v;
            check(std::fabs(vsum - 1.0) < 1e-12, "control volumes tile the unit cube");
            // the same mesh through the tetrahedra constructor
            vfvm_mesh* m2 = nullptr;
            check(vfvm_mesh_from_tets(info.num_pts, info.num_nds, pos.data(), info.num_tets, tets.data(), info.num_bnd, bptr.data(),
                                      bnid.data(), &m2) == VFVM_OK,
                  "vfvm_mesh_from_tets");
            if (m2) {
                std::vector<double> vol2(N);
                check(vfvm_mesh_get(m2, VFVM_MESH_VOL, vol2.data()) == VFVM_OK && vol2 == vol, "from_tets dual equals lattice dual");
                vfvm_mesh_destroy(m2);
            }
            // the device builders are absent here and report it
            check(vfvm_mesh_build_grad_coeff_dev(m, 0) != VFVM_OK, "device builder reports its absence");
            // ILU(0) ordering of the scalar pattern and of a two-unknown block pattern
            std::vector<int32_t> mrow, mcol;
            scalar_pattern(info, rp, ngh, mrow, mcol);
            check_ilu_order(mrow, mcol, pos, N);
            std::vector<int32_t> brow(2 * N + 1, 0), bcol;
            for (int b = 0; b < 2; ++b)
                for (int64_t v = 0; v < N; ++v) {
                    for (int32_t k = mrow[v]; k < mrow[v + 1]; ++k) bcol.push_back((int32_t)(b * N + mcol[k]));
                    if (b == 0) bcol.push_back((int32_t)(N + v));  // coupling into the second block
                    brow[b * N + v + 1] = (int32_t)bcol.size();
                }
            check_ilu_order(brow, bcol, pos, N);
            // output in both formats, with a value whose fixed-point expansion is 312 characters
            std::vector<double> f(N);
            for (int64_t v = 0; v < N; ++v) f[v] = std::sin(3.0 * pos[3 * v]) * std::pow(10.0, (double)(v % 7) * 40.0 - 120.0);
            f[0] = -1e300;
            f[1] = NAN;
            check(vfvm_write_csv(m, -1, 1, f.data(), "/tmp/vfvm_host_check", "csv", n, 9) == VFVM_OK, "vfvm_write_csv");
            check(vfvm_write_vtu(m, -1, 1, f.data(), "/tmp/vfvm_host_check", "vtu", n, 0) == VFVM_OK, "vfvm_write_vtu");
            check(vfvm_write_csv(m, 2, 1, f.data(), "/tmp/vfvm_host_check_b", "csv", n, 0) == VFVM_OK, "boundary csv");
            vfvm_mesh_destroy(m);
        }
    }
    // invalid input: a tetrahedron whose vertex ids run out of range, a point on the hull
    const double tp[12] = {1, 1, 1, 1, -1, -1, -1, 1, -1, -1, -1, 1};
    const int32_t bad[4] = {0, 1, 2, 7};
    const int32_t tt[4] = {0, 1, 2, 3};
    const int64_t bp[2] = {0, 3};
    const int32_t bn[3] = {0, 1, 2};
    vfvm_mesh* m = nullptr;
    check(vfvm_mesh_from_tets(0, 4, tp, 1, bad, 1, bp, bn, &m) == VFVM_ERR_INVALID_ARGUMENT, "out-of-range tetrahedron rejected");
    check(vfvm_mesh_from_tets(1, 3, tp, 1, tt, 1, bp, bn, &m) == VFVM_ERR_TETRA_POINTS_ON_BOUNDARY, "point on the hull rejected");
    if (fails) return 1;
    std::printf("host check ok\n");
    return 0;
}
