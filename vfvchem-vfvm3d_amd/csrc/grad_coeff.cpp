// Pressure-gradient coefficients of the flow operators (restates src/base/geom_tet.rs:295-347).
//
// For vertex v, the divergence-theorem gradient with skew-corrected face values
//   int_V grad(p) dV = sum_faces p_f n A + sum_subfaces p_b n A = sum_m (c_x, c_y, c_z)_m p_m
// with p_f = (p_v + p_n)/2 + (grad p_v + grad p_n)/2 . skew, grad p_s = sum_loc w (p_m - p_s)
// (fully implicit skew correction) and p_b = w_v p_v + w_j p_j + w_k p_k.  The terms are
// pushed in the reference's order, sorted by vertex with a stable sort and merged in that
// order, so each coefficient is the same sum of the same terms in the same order.
#include <algorithm>
#include <thread>
#include <vector>

#include "mesh.hpp"
#include "status.hpp"

namespace vfvm {

namespace {

struct Term {
    int32_t vid;
    double cx, cy, cz;
};

void coeff_of(const Mesh& m, int64_t v, std::vector<Term>& t) {
    const int64_t E = m.num_edges(), S = m.num_subfaces();
    t.clear();
    for (int32_t e = m.row_ptr[v]; e < m.row_ptr[v + 1]; ++e) {
        const int32_t n = m.ngh[e];
        const double area = m.area[e];
        const double nx = m.norm[e], ny = m.norm[E + e], nz = m.norm[2 * E + e];
        const double sx = m.skew[e], sy = m.skew[E + e], sz = m.skew[2 * E + e];
        // arithmetic mean
        t.push_back({(int32_t)v, 0.5 * area * nx, 0.5 * area * ny, 0.5 * area * nz});
        t.push_back({n, 0.5 * area * nx, 0.5 * area * ny, 0.5 * area * nz});
        // skewness correction: (p_m - p_s) terms for s = v, n
        for (int32_t s : {(int32_t)v, n}) {
            for (int32_t q = m.row_ptr[s]; q < m.row_ptr[s + 1]; ++q) {
                const double wx = m.gw[q], wy = m.gw[E + q], wz = m.gw[2 * E + q];
                const double w_skew = 0.5 * (wx * sx + wy * sy + wz * sz) * area;
                t.push_back({m.ngh[q], w_skew * nx, w_skew * ny, w_skew * nz});
                t.push_back({s, -w_skew * nx, -w_skew * ny, -w_skew * nz});
            }
        }
    }
    // boundary sub-faces: linear interpolation to the sub-face centroid
    for (int32_t b = m.bsf_ptr[v]; b < m.bsf_ptr[v + 1]; ++b) {
        const double area = m.bsf_area[b];
        const double nx = m.bsf_norm[b], ny = m.bsf_norm[S + b], nz = m.bsf_norm[2 * S + b];
        const double wv = m.bsf_wgt[b], wj = m.bsf_wgt[S + b], wk = m.bsf_wgt[2 * S + b];
        t.push_back({(int32_t)v, wv * area * nx, wv * area * ny, wv * area * nz});
        t.push_back({m.bsf_vj[b], wj * area * nx, wj * area * ny, wj * area * nz});
        t.push_back({m.bsf_vk[b], wk * area * nx, wk * area * ny, wk * area * nz});
    }
    // merge repeated vertices: stable sort, then sum in push order
    std::stable_sort(t.begin(), t.end(), [](const Term& a, const Term& b) { return a.vid < b.vid; });
    size_t w = 0;
    for (size_t i = 0; i < t.size(); ++i) {
        if (w > 0 && t[w - 1].vid == t[i].vid) {
            t[w - 1].cx += t[i].cx;
            t[w - 1].cy += t[i].cy;
            t[w - 1].cz += t[i].cz;
        } else {
            t[w++] = t[i];
        }
    }
    t.resize(w);
}

}  // namespace

int build_grad_coeff(Mesh& m, std::string& err) {
    if (!m.has_dual) {
        err = "grad_coeff: mesh without dual";
        return VFVM_ERR_INVALID_ARGUMENT;
    }
    if (m.has_grad_coeff) return VFVM_OK;
    const int64_t N = m.num_tot();
    unsigned nth = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::vector<std::vector<int32_t>> cnt(nth);
    std::vector<std::vector<Term>> parts(nth);
    auto work = [&](unsigned tid) {
        const int64_t a = N * tid / nth, b = N * (tid + 1) / nth;
        std::vector<Term> t;
        cnt[tid].resize(b - a);
        for (int64_t v = a; v < b; ++v) {
            coeff_of(m, v, t);
            cnt[tid][v - a] = (int32_t)t.size();
            parts[tid].insert(parts[tid].end(), t.begin(), t.end());
        }
    };
    std::vector<std::thread> th;
    for (unsigned k = 1; k < nth; ++k) th.emplace_back(work, k);
    work(0);
    for (auto& x : th) x.join();
    m.gc_ptr.assign(N + 1, 0);
    int64_t total = 0;
    for (unsigned k = 0; k < nth; ++k) {
        const int64_t a = N * k / nth;
        for (size_t i = 0; i < cnt[k].size(); ++i) {
            total += cnt[k][i];
            if (total > INT32_MAX) {
                err = "grad_coeff: more than 2^31 coefficients";
                return VFVM_ERR_INVALID_ARGUMENT;
            }
            m.gc_ptr[a + i + 1] = (int32_t)total;
        }
    }
    m.gc_vid.resize(total);
    m.gc_c.resize(3 * total);
    int64_t pos = 0;
    for (unsigned k = 0; k < nth; ++k) {
        for (const Term& t : parts[k]) {
            m.gc_vid[pos] = t.vid;
            m.gc_c[pos] = t.cx;
            m.gc_c[total + pos] = t.cy;
            m.gc_c[2 * total + pos] = t.cz;
            ++pos;
        }
        parts[k].clear();
        parts[k].shrink_to_fit();
    }
    m.has_grad_coeff = true;
    return VFVM_OK;
}

}  // namespace vfvm
