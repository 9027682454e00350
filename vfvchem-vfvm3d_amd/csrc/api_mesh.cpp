// C ABI: mesh generation and access (host only).
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "mesh.hpp"

using namespace vfvm;

struct vfvm_mesh {
    Mesh m;
};

// the dual on the host (dual_device < 0) or on that GPU
static int make_dual(Mesh& m, int32_t dual_device, std::string& err) {
    return dual_device < 0 ? build_dual(m, 0, err) : build_dual_device(m, dual_device, err);
}

static int mesh_lattice(int32_t n, uint64_t seed, double jitter, int32_t dual_device, vfvm_mesh** out) {
    if (!out) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_lattice: out is NULL");
    *out = nullptr;
    auto* h = new vfvm_mesh();
    std::string err;
    int st = lattice_mesh(n, seed, jitter, h->m, err);
    if (st == VFVM_OK) st = make_dual(h->m, dual_device, err);
    if (st != VFVM_OK) { delete h; return set_error(st, err); }
    *out = h;
    return clear_error();
}

extern "C" int vfvm_mesh_lattice(int32_t n, uint64_t seed, double jitter, vfvm_mesh** out) {
    return mesh_lattice(n, seed, jitter, -1, out);
}
extern "C" int vfvm_mesh_lattice_dev(int32_t n, uint64_t seed, double jitter, int32_t device, vfvm_mesh** out) {
    if (device < 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_lattice_dev: negative device");
    return mesh_lattice(n, seed, jitter, device, out);
}

extern "C" int vfvm_mesh_lattice_box(int32_t n, uint64_t seed, double jitter, const int32_t* lo, const int32_t* hi,
                                     int32_t device, vfvm_mesh** out) {
    if (!out || !lo || !hi) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_lattice_box: NULL argument");
    *out = nullptr;
    auto* h = new vfvm_mesh();
    std::string err;
    int st = lattice_box_mesh(n, seed, jitter, lo, hi, h->m, err);
    if (st == VFVM_OK) st = make_dual(h->m, device, err);
    if (st != VFVM_OK) { delete h; return set_error(st, err); }
    *out = h;
    return clear_error();
}

extern "C" int vfvm_mesh_join(int32_t num_dom, const vfvm_mesh* const* doms, int32_t num_itf, const int32_t* itf,
                              vfvm_mesh** out) {
    if (!out || !doms || num_dom < 1 || num_itf < 0 || (num_itf > 0 && !itf))
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_join: invalid argument");
    *out = nullptr;
    std::vector<const Mesh*> parts;
    for (int32_t d = 0; d < num_dom; ++d) parts.push_back(doms[d] ? &doms[d]->m : nullptr);
    std::vector<JoinItf> fs;
    for (int32_t k = 0; k < num_itf; ++k) fs.push_back({itf[4 * k], itf[4 * k + 1], itf[4 * k + 2], itf[4 * k + 3]});
    auto* h = new vfvm_mesh();
    std::string err;
    int st = join_meshes(parts, fs, h->m, err);
    if (st != VFVM_OK) { delete h; return set_error(st, err); }
    *out = h;
    return clear_error();
}

static int mesh_from_tets(int64_t num_pts, int64_t num_nds, const double* pos, int64_t num_tets, const int32_t* tets,
                          int32_t num_bnd, const int64_t* bnd_ptr, const int32_t* bnd_nid, int32_t dual_device,
                          vfvm_mesh** out) {
    if (!out || !pos || !tets || num_pts < 0 || num_nds < 0 || num_tets <= 0 || num_bnd < 0 || (num_bnd > 0 && (!bnd_ptr || !bnd_nid)))
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_from_tets: invalid argument");
    *out = nullptr;
    const int64_t N = num_pts + num_nds;
    if (N > INT32_MAX) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_from_tets: too many vertices");
    auto* h = new vfvm_mesh();
    Mesh& m = h->m;
    m.num_pts = num_pts;
    m.num_nds = num_nds;
    m.pos.assign(pos, pos + 3 * N);
    m.tets.assign(tets, tets + 4 * num_tets);
    // orient every tetrahedron as the reference does (p3 below the counterclockwise face p0 p1 p2)
    for (int64_t t = 0; t < num_tets; ++t) {
        int32_t* q = &m.tets[4 * t];
        for (int k = 0; k < 4; ++k)
            if (q[k] < 0 || q[k] >= N) { delete h; return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_from_tets: vertex id out of range"); }
        const double *a = &m.pos[3 * (int64_t)q[0]], *b = &m.pos[3 * (int64_t)q[1]], *c = &m.pos[3 * (int64_t)q[2]], *d = &m.pos[3 * (int64_t)q[3]];
        double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]}, w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
        double det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0]);
        if (det == 0.0) { delete h; return set_error(VFVM_ERR_TETRA_DEGENERATE_INPUT, "vfvm_mesh_from_tets: zero-volume tetrahedron"); }
        if (det > 0.0) std::swap(q[1], q[2]);
    }
    m.bnd_ptr.assign(num_bnd + 1, 0);
    m.nd_bnd.assign(num_nds, -1);
    for (int32_t b = 0; b < num_bnd; ++b) {
        m.bnd_name.push_back("b" + std::to_string(b));
        m.bnd_ptr[b + 1] = bnd_ptr[b + 1] - bnd_ptr[0];
        for (int64_t k = bnd_ptr[b]; k < bnd_ptr[b + 1]; ++k) {
            int32_t nid = bnd_nid[k];
            if (nid < 0 || nid >= num_nds || m.nd_bnd[nid] != -1) { delete h; return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_from_tets: node id out of range or on two boundaries"); }
            m.nd_bnd[nid] = b;
            m.bnd_nid.push_back(nid);
        }
    }
    std::string err;
    int st = make_dual(m, dual_device, err);
    if (st != VFVM_OK) { delete h; return set_error(st, err); }
    *out = h;
    return clear_error();
}

extern "C" int vfvm_mesh_from_tets(int64_t num_pts, int64_t num_nds, const double* pos,
                                   int64_t num_tets, const int32_t* tets,
                                   int32_t num_bnd, const int64_t* bnd_ptr, const int32_t* bnd_nid,
                                   vfvm_mesh** out) {
    return mesh_from_tets(num_pts, num_nds, pos, num_tets, tets, num_bnd, bnd_ptr, bnd_nid, -1, out);
}
extern "C" int vfvm_mesh_from_tets_dev(int64_t num_pts, int64_t num_nds, const double* pos,
                                       int64_t num_tets, const int32_t* tets,
                                       int32_t num_bnd, const int64_t* bnd_ptr, const int32_t* bnd_nid,
                                       int32_t device, vfvm_mesh** out) {
    if (device < 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_from_tets_dev: negative device");
    return mesh_from_tets(num_pts, num_nds, pos, num_tets, tets, num_bnd, bnd_ptr, bnd_nid, device, out);
}

extern "C" int vfvm_mesh_info_get(const vfvm_mesh* mesh, vfvm_mesh_info* info) {
    if (!mesh || !info) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_info_get: NULL argument");
    const Mesh& m = mesh->m;
    info->num_pts = m.num_pts;
    info->num_nds = m.num_nds;
    info->num_tets = (int64_t)m.tets.size() / 4;
    info->num_edges = m.num_edges();
    info->num_subfaces = m.num_subfaces();
    info->num_bnd = (int32_t)m.bnd_name.size();
    int32_t md = 0;
    for (int64_t v = 0; v + 1 < (int64_t)m.row_ptr.size(); ++v) md = std::max(md, m.row_ptr[v + 1] - m.row_ptr[v]);
    info->max_degree = md;
    info->num_grad_coeff = m.has_grad_coeff ? (int64_t)m.gc_vid.size() : 0;
    info->num_dom = m.num_dom();
    info->num_itf = m.num_itf();
    info->num_itf_nodes = (int64_t)m.itf_va.size();
    return clear_error();
}

extern "C" int vfvm_mesh_build_grad_coeff(vfvm_mesh* mesh) {
    if (!mesh) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_build_grad_coeff: NULL mesh");
    if (mesh->m.num_dom() > 1 || mesh->m.num_itf() > 0)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "pressure-gradient coefficients of a joined mesh (flow across interfaces) are outside the built path");
    std::string err;
    int st = build_grad_coeff(mesh->m, err);
    if (st != VFVM_OK) return set_error(st, err);
    return clear_error();
}
extern "C" int vfvm_mesh_build_grad_coeff_dev(vfvm_mesh* mesh, int32_t device) {
    if (!mesh || device < 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_build_grad_coeff_dev: NULL mesh or negative device");
    if (mesh->m.num_dom() > 1 || mesh->m.num_itf() > 0)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "pressure-gradient coefficients of a joined mesh (flow across interfaces) are outside the built path");
    std::string err;
    int st = build_grad_coeff_device(mesh->m, device, err);
    if (st != VFVM_OK) return set_error(st, err);
    return clear_error();
}

template <class T>
static int copy_out(const std::vector<T>& v, void* dst) {
    if (!v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
    return clear_error();
}

extern "C" int vfvm_mesh_get(const vfvm_mesh* mesh, int32_t id, void* dst) {
    if (!mesh || !dst) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_get: NULL argument");
    const Mesh& m = mesh->m;
    switch (id) {
        case VFVM_MESH_POS: return copy_out(m.pos, dst);
        case VFVM_MESH_TETS: return copy_out(m.tets, dst);
        case VFVM_MESH_BND_PTR: return copy_out(m.bnd_ptr, dst);
        case VFVM_MESH_BND_NID: return copy_out(m.bnd_nid, dst);
        case VFVM_MESH_ROW_PTR: return copy_out(m.row_ptr, dst);
        case VFVM_MESH_NGH: return copy_out(m.ngh, dst);
        case VFVM_MESH_DIST: return copy_out(m.dist, dst);
        case VFVM_MESH_AREA: return copy_out(m.area, dst);
        case VFVM_MESH_NORM: return copy_out(m.norm, dst);
        case VFVM_MESH_SKEW: return copy_out(m.skew, dst);
        case VFVM_MESH_GRAD_WGT: return copy_out(m.gw, dst);
        case VFVM_MESH_VOL: return copy_out(m.vol, dst);
        case VFVM_MESH_CSKEW: return copy_out(m.cskew, dst);
        case VFVM_MESH_BSF_PTR: return copy_out(m.bsf_ptr, dst);
        case VFVM_MESH_BSF_VJ: return copy_out(m.bsf_vj, dst);
        case VFVM_MESH_BSF_VK: return copy_out(m.bsf_vk, dst);
        case VFVM_MESH_BSF_AREA: return copy_out(m.bsf_area, dst);
        case VFVM_MESH_BSF_NORM: return copy_out(m.bsf_norm, dst);
        case VFVM_MESH_BSF_WGT: return copy_out(m.bsf_wgt, dst);
        case VFVM_MESH_GC_PTR: return copy_out(m.gc_ptr, dst);
        case VFVM_MESH_GC_VID: return copy_out(m.gc_vid, dst);
        case VFVM_MESH_GC_C: return copy_out(m.gc_c, dst);
        case VFVM_MESH_DOM_PTR: return copy_out(m.dom_ptr.empty() ? std::vector<int64_t>{0, m.num_tot()} : m.dom_ptr, dst);
        case VFVM_MESH_DOM_NPTS: return copy_out(m.dom_npts.empty() ? std::vector<int64_t>{m.num_pts} : m.dom_npts, dst);
        case VFVM_MESH_ITF_PTR: return copy_out(m.itf_ptr.empty() ? std::vector<int64_t>{0} : m.itf_ptr, dst);
        case VFVM_MESH_ITF_VA: return copy_out(m.itf_va, dst);
        case VFVM_MESH_ITF_VB: return copy_out(m.itf_vb, dst);
        case VFVM_MESH_ITF_DOM: return copy_out(m.itf_dom, dst);
        default: return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_get: unknown array id");
    }
}

extern "C" void vfvm_mesh_destroy(vfvm_mesh* mesh) { delete mesh; }

const vfvm::Mesh* vfvm_mesh_ptr(const vfvm_mesh* mesh) { return mesh ? &mesh->m : nullptr; }
