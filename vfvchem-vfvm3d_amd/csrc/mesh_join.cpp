// Multi-domain meshes: several domains, each with its own vertex set and Voronoi dual, joined
// by interfaces whose nodes belong to both domains (Interface, src/base/geom_itf.rs:6-49;
// one Tetrahedron per domain, src/base/vars.rs:7-34).
//
// The joined mesh numbers domain d's vertices [off_d, off_d + N_d) in the domain's own order
// (its points, then its nodes), so a scalar unknown over the joined vertices is the
// reference's sequence of per-domain unknown scalars with their blocks started in domain
// order (SteadyBase::initial_start).  Each domain keeps its own dual: the neighbour CSR is
// block diagonal, and an interface node has one copy per side with the faces of that side's
// half of its control volume.  Boundary groups named as interface sides become interfaces;
// the others stay boundaries, numbered domain by domain.
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "mesh.hpp"
#include "status.hpp"

namespace vfvm {

namespace {

template <class T>
void append_shift(std::vector<T>& dst, const std::vector<T>& src, T shift) {
    for (const T& x : src) dst.push_back(x + shift);
}

// SoA arrays of c components, each of length len_d per domain: component-wise concatenation
void append_soa(std::vector<double>& dst, const std::vector<const std::vector<double>*>& parts,
                const std::vector<int64_t>& len, int comps) {
    for (int c = 0; c < comps; ++c)
        for (size_t d = 0; d < parts.size(); ++d) {
            const std::vector<double>& s = *parts[d];
            dst.insert(dst.end(), s.begin() + c * len[d], s.begin() + (c + 1) * len[d]);
        }
}

}  // namespace

int join_meshes(const std::vector<const Mesh*>& doms, const std::vector<JoinItf>& itfs, Mesh& out, std::string& err) {
    const int32_t nd = (int32_t)doms.size();
    if (nd < 1) { err = "join: no domain"; return VFVM_ERR_INVALID_ARGUMENT; }
    for (int32_t d = 0; d < nd; ++d) {
        if (!doms[d] || !doms[d]->has_dual) { err = "join: domain " + std::to_string(d) + " has no dual"; return VFVM_ERR_INVALID_ARGUMENT; }
        if (doms[d]->num_dom() != 1) { err = "join: domain " + std::to_string(d) + " is itself a joined mesh"; return VFVM_ERR_INVALID_ARGUMENT; }
    }
    // boundaries used as interface sides
    std::vector<std::vector<int>> side_of(nd);  // per domain, per boundary: -1 or interface index
    for (int32_t d = 0; d < nd; ++d) side_of[d].assign(doms[d]->bnd_name.size(), -1);
    for (size_t k = 0; k < itfs.size(); ++k) {
        const JoinItf& f = itfs[k];
        if (f.dom_a < 0 || f.dom_a >= nd || f.dom_b < 0 || f.dom_b >= nd || f.dom_a == f.dom_b) {
            err = "join: interface " + std::to_string(k) + ": domain ids out of range or equal";
            return VFVM_ERR_INVALID_ARGUMENT;
        }
        if (f.bnd_a < 0 || f.bnd_a >= (int32_t)side_of[f.dom_a].size() || f.bnd_b < 0 || f.bnd_b >= (int32_t)side_of[f.dom_b].size()) {
            err = "join: interface " + std::to_string(k) + ": boundary id out of range";
            return VFVM_ERR_INVALID_ARGUMENT;
        }
        if (side_of[f.dom_a][f.bnd_a] >= 0 || side_of[f.dom_b][f.bnd_b] >= 0) {
            err = "join: interface " + std::to_string(k) + ": a boundary is already an interface side";
            return VFVM_ERR_INVALID_ARGUMENT;
        }
        side_of[f.dom_a][f.bnd_a] = (int)k;
        side_of[f.dom_b][f.bnd_b] = (int)k;
    }

    out = Mesh();
    Mesh& m = out;
    std::vector<int64_t> off(nd + 1, 0), eoff(nd + 1, 0), soff(nd + 1, 0), toff(nd + 1, 0);
    for (int32_t d = 0; d < nd; ++d) {
        off[d + 1] = off[d] + doms[d]->num_tot();
        eoff[d + 1] = eoff[d] + doms[d]->num_edges();
        soff[d + 1] = soff[d] + doms[d]->num_subfaces();
        toff[d + 1] = toff[d] + (int64_t)doms[d]->tets.size() / 4;
    }
    const int64_t N = off[nd];
    if (N > INT32_MAX || eoff[nd] > INT32_MAX || soff[nd] > INT32_MAX) {
        err = "join: more than 2^31 vertices, edges or sub-faces";
        return VFVM_ERR_INVALID_ARGUMENT;
    }
    m.num_pts = 0;
    m.num_nds = N;
    m.dom_ptr = off;
    m.dom_npts.resize(nd);
    std::vector<int64_t> lenN(nd), lenE(nd), lenS(nd);
    std::vector<const std::vector<double>*> P;
    for (int32_t d = 0; d < nd; ++d) {
        const Mesh& s = *doms[d];
        m.dom_npts[d] = s.num_pts;
        lenN[d] = s.num_tot();
        lenE[d] = s.num_edges();
        lenS[d] = s.num_subfaces();
        m.pos.insert(m.pos.end(), s.pos.begin(), s.pos.end());
        append_shift(m.tets, s.tets, (int32_t)off[d]);
        m.ngh.reserve(eoff[nd]);
        append_shift(m.ngh, s.ngh, (int32_t)off[d]);
        m.dist.insert(m.dist.end(), s.dist.begin(), s.dist.end());
        m.area.insert(m.area.end(), s.area.begin(), s.area.end());
        m.vol.insert(m.vol.end(), s.vol.begin(), s.vol.end());
        m.bnd_sum.insert(m.bnd_sum.end(), s.bnd_sum.begin(), s.bnd_sum.end());
        append_shift(m.bsf_vj, s.bsf_vj, (int32_t)off[d]);
        append_shift(m.bsf_vk, s.bsf_vk, (int32_t)off[d]);
        m.bsf_area.insert(m.bsf_area.end(), s.bsf_area.begin(), s.bsf_area.end());
    }
    m.row_ptr.assign(1, 0);
    m.bsf_ptr.assign(1, 0);
    for (int32_t d = 0; d < nd; ++d) {
        const Mesh& s = *doms[d];
        for (int64_t v = 0; v < s.num_tot(); ++v) {
            m.row_ptr.push_back((int32_t)(s.row_ptr[v + 1] + eoff[d]));
            m.bsf_ptr.push_back((int32_t)(s.bsf_ptr[v + 1] + soff[d]));
        }
    }
    auto soa = [&](std::vector<double>& dst, std::vector<double> Mesh::*field, const std::vector<int64_t>& len, int comps) {
        std::vector<const std::vector<double>*> parts;
        for (int32_t d = 0; d < nd; ++d) parts.push_back(&(doms[d]->*field));
        append_soa(dst, parts, len, comps);
    };
    soa(m.norm, &Mesh::norm, lenE, 3);
    soa(m.skew, &Mesh::skew, lenE, 3);
    soa(m.gw, &Mesh::gw, lenE, 3);
    soa(m.cskew, &Mesh::cskew, lenN, 3);
    soa(m.bsf_norm, &Mesh::bsf_norm, lenS, 3);
    soa(m.bsf_wgt, &Mesh::bsf_wgt, lenS, 3);
    m.has_dual = true;

    // boundaries that are not interface sides, domain by domain (node ids = vertex ids)
    m.nd_bnd.assign(N, -1);
    m.bnd_ptr.assign(1, 0);
    for (int32_t d = 0; d < nd; ++d) {
        const Mesh& s = *doms[d];
        for (size_t b = 0; b < s.bnd_name.size(); ++b) {
            if (side_of[d][b] >= 0) continue;
            const int32_t id = (int32_t)m.bnd_name.size();
            m.bnd_name.push_back("d" + std::to_string(d) + "." + s.bnd_name[b]);
            for (int64_t k = s.bnd_ptr[b]; k < s.bnd_ptr[b + 1]; ++k) {
                const int32_t v = (int32_t)(off[d] + s.num_pts + s.bnd_nid[k]);
                m.bnd_nid.push_back(v);
                m.nd_bnd[v] = id;
            }
            m.bnd_ptr.push_back((int64_t)m.bnd_nid.size());
        }
    }
    // interfaces: side a's nodes in its group order, each paired with the side-b node at the
    // same coordinates
    m.itf_ptr.assign(1, 0);
    for (size_t k = 0; k < itfs.size(); ++k) {
        const JoinItf& f = itfs[k];
        const Mesh& A = *doms[f.dom_a];
        const Mesh& B = *doms[f.dom_b];
        const int64_t na = A.bnd_ptr[f.bnd_a + 1] - A.bnd_ptr[f.bnd_a];
        const int64_t nb = B.bnd_ptr[f.bnd_b + 1] - B.bnd_ptr[f.bnd_b];
        if (na != nb) {
            err = "join: interface " + std::to_string(k) + ": the two sides hold " + std::to_string(na) + " and " +
                  std::to_string(nb) + " nodes";
            return VFVM_ERR_INVALID_ARGUMENT;
        }
        std::map<std::tuple<double, double, double>, int32_t> at;
        for (int64_t q = B.bnd_ptr[f.bnd_b]; q < B.bnd_ptr[f.bnd_b + 1]; ++q) {
            const int64_t v = B.num_pts + B.bnd_nid[q];
            at[{B.pos[3 * v], B.pos[3 * v + 1], B.pos[3 * v + 2]}] = (int32_t)(off[f.dom_b] + v);
        }
        for (int64_t q = A.bnd_ptr[f.bnd_a]; q < A.bnd_ptr[f.bnd_a + 1]; ++q) {
            const int64_t v = A.num_pts + A.bnd_nid[q];
            auto it = at.find({A.pos[3 * v], A.pos[3 * v + 1], A.pos[3 * v + 2]});
            if (it == at.end()) {
                err = "join: interface " + std::to_string(k) + ": node " + std::to_string(v) + " of side a has no copy on side b";
                return VFVM_ERR_INVALID_ARGUMENT;
            }
            m.itf_va.push_back((int32_t)(off[f.dom_a] + v));
            m.itf_vb.push_back(it->second);
        }
        m.itf_ptr.push_back((int64_t)m.itf_va.size());
        m.itf_dom.push_back(f.dom_a);
        m.itf_dom.push_back(f.dom_b);
    }
    return VFVM_OK;
}

}  // namespace vfvm
