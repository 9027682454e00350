// Host-side mesh: jittered-lattice vertex set, exact Delaunay tetrahedralisation
// and the Voronoi dual arrays that the device kernels read.
//
// Vertex numbering follows the reference: points (interior vertices) first, then
// nodes (boundary vertices), vid = pid for points and vid = num_pts + nid for nodes
// (src/readme.txt:58-59, src/base/geom_tet.rs:11-12).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace vfvm {

struct Mesh {
    // vertices
    int64_t num_pts = 0;            // interior points
    int64_t num_nds = 0;            // boundary nodes
    std::vector<double> pos;        // 3 * num_tot, AoS (x, y, z)
    int64_t num_tot() const { return num_pts + num_nds; }

    // boundaries: node ids (nid, local to the node block) per boundary, ascending
    std::vector<std::string> bnd_name;
    std::vector<int64_t> bnd_ptr;   // num_bnd + 1
    std::vector<int32_t> bnd_nid;   // concatenated nids
    std::vector<int32_t> nd_bnd;    // nid -> boundary id (each node on exactly one boundary)

    // tetrahedra, reference orientation (p3 below the counterclockwise face p0 p1 p2)
    std::vector<int32_t> tets;      // 4 * num_tets

    // ---- Voronoi dual (filled by build_dual; restates src/base/geom_tet.rs:58-350) ----
    bool has_dual = false;
    std::vector<int32_t> row_ptr;   // num_tot + 1; neighbour CSR (directed edges)
    std::vector<int32_t> ngh;       // E; ascending per row (geom_tet.rs:174-196)
    std::vector<double> dist;       // E; edge length
    std::vector<double> area;       // E; signed Voronoi face area
    std::vector<double> norm;       // 3E SoA: nx[E], ny[E], nz[E]; outward unit normal
    std::vector<double> skew;       // 3E SoA; face centre minus edge midpoint
    std::vector<double> gw;         // 3E SoA; least-squares gradient weights
    std::vector<double> vol;        // N; control-volume volume
    std::vector<double> cskew;      // 3N SoA; control-volume centroid minus vertex
    // boundary sub-faces (geom_tet.rs:208-259)
    std::vector<int32_t> bsf_ptr;   // N + 1
    std::vector<int32_t> bsf_vj, bsf_vk;  // other two vids of the surface triangle
    std::vector<double> bsf_area;   // signed sub-face area
    std::vector<double> bsf_norm;   // 3 * S SoA
    std::vector<double> bsf_wgt;    // 3 * S SoA (w_v, w_j, w_k)
    std::vector<double> bnd_sum;    // N; boundary CV area
    // pressure-gradient coefficients (geom_tet.rs:295-347), built on demand by build_grad_coeff:
    // CSR over vertices of (vid_m, c_x, c_y, c_z), vid_m ascending (the 2-ring of v)
    bool has_grad_coeff = false;
    std::vector<int32_t> gc_ptr;    // N + 1
    std::vector<int32_t> gc_vid;    // G
    std::vector<double> gc_c;       // 3G SoA

    // multi-domain meshes (join_meshes): domain d holds the vertices [dom_ptr[d], dom_ptr[d+1]),
    // the first dom_npts[d] of them its points, then its nodes; num_pts is 0 and node ids are
    // vertex ids.  Interface k pairs the two copies itf_va[q] (side a) and itf_vb[q] (side b)
    // of one node, q in [itf_ptr[k], itf_ptr[k+1]), in the order of side a's node group.
    // Empty for a single-domain mesh.
    std::vector<int64_t> dom_ptr, dom_npts;
    std::vector<int64_t> itf_ptr;
    std::vector<int32_t> itf_va, itf_vb;
    std::vector<int32_t> itf_dom;   // 2 per interface: domain of side a, of side b
    int32_t num_dom() const { return dom_ptr.empty() ? 1 : (int32_t)dom_ptr.size() - 1; }
    int32_t num_itf() const { return itf_ptr.empty() ? 0 : (int32_t)itf_ptr.size() - 1; }

    int64_t num_edges() const { return (int64_t)ngh.size(); }
    int64_t num_subfaces() const { return (int64_t)bsf_vj.size(); }
};

// Status codes shared with the C ABI (mirrors of VFVChemError variants on this path).
int lattice_mesh(int n, uint64_t seed, double jitter, Mesh& m, std::string& err);
int lattice_box_mesh(int n, uint64_t seed, double jitter, const int32_t lo[3], const int32_t hi[3], Mesh& m,
                     std::string& err);
// Interface of a join: boundary bnd_a of domain dom_a and boundary bnd_b of domain dom_b hold
// the same nodes (equal coordinates).
struct JoinItf {
    int32_t dom_a, bnd_a, dom_b, bnd_b;
};
// One mesh of several domains whose duals are built (mesh_join.cpp).
int join_meshes(const std::vector<const Mesh*>& doms, const std::vector<JoinItf>& itfs, Mesh& out, std::string& err);
int delaunay_int(const std::vector<int64_t>& ixyz, const std::vector<int32_t>& order,
                 std::vector<int32_t>& tets_out, std::string& err);
int build_dual(Mesh& m, int nthreads, std::string& err);
// The same dual built on GPU `device` (bitwise equal to build_dual; dual_gpu.hip).
int build_dual_device(Mesh& m, int device, std::string& err);
// Pieces of build_dual shared with the device build.
std::vector<int32_t> dual_surface_triangles(const Mesh& m, const std::vector<uint8_t>& is_surf);
int dual_check_surface(const Mesh& m, const std::vector<int32_t>& surf, std::string& err);
void dual_subfaces(Mesh& m, const std::vector<int32_t>& surf);
int build_grad_coeff(Mesh& m, std::string& err);
// The same coefficients built on GPU `device`, bitwise equal (dual_gpu.hip).
int build_grad_coeff_device(Mesh& m, int device, std::string& err);

}  // namespace vfvm
