// Assembly-side kernels: least-squares gradients, prescribed-velocity face flux,
// and the row-gather operator assembly into the fixed CSR pattern.
//
// Every row is produced by one thread that walks its directed edges in CSR
// order, so each matrix entry and each b entry is summed in the order the
// reference pushes its triplets (operator order, then neighbour order), with
// no atomics.  The file is compiled with -ffp-contract=off so that products
// and sums round exactly as the reference's f64 expressions do.
#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "xcd.hpp"

namespace vfvm {

namespace {

constexpr int BLK = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

// ScalarDomain::update_grad (src/base/scl_dom.rs:146-164): grad u_c = sum_loc w_loc (u_n - u_c)
// accumulated in loc order.  LPR lanes per vertex load the edge weights of a row
// contiguously; the sum walks the lanes' terms with shuffles in loc order (bit-identical to
// a sequential walk).  Rows of degree > 2 LPR take the sequential walk in their first lane.
template <int LPR>
__global__ void __launch_bounds__(BLK) k_grad(int64_t N, int64_t E, const int32_t* __restrict__ rp,
                                              const int32_t* __restrict__ ngh, const double* __restrict__ gw,
                                              const double* __restrict__ val, double* __restrict__ grad) {
    const int64_t tid = (int64_t)xcd_block() * BLK + threadIdx.x;
    const int64_t v = tid / LPR;
    const int lane = (int)(tid % LPR);
    if (v >= N) return;
    const double uc = val[v];
    const int32_t es = rp[v], ee = rp[v + 1], deg = ee - es;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (deg <= 2 * LPR) {
        double t[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        for (int h = 0; h < 2; ++h) {
            const int32_t e = es + h * LPR + lane;
            if (e < ee) {
                const double d = val[ngh[e]] - uc;
                t[h][0] = gw[e] * d;
                t[h][1] = gw[E + e] * d;
                t[h][2] = gw[2 * E + e] * d;
            }
        }
        for (int k = 0; k < deg; ++k) {
            const int h = k / LPR, src = k % LPR;
            gx += __shfl(h ? t[1][0] : t[0][0], src, LPR);
            gy += __shfl(h ? t[1][1] : t[0][1], src, LPR);
            gz += __shfl(h ? t[1][2] : t[0][2], src, LPR);
        }
    } else if (lane == 0) {
        for (int32_t e = es; e < ee; ++e) {
            const double d = val[ngh[e]] - uc;
            gx += gw[e] * d;
            gy += gw[E + e] * d;
            gz += gw[2 * E + e] * d;
        }
    }
    if (lane == 0) {
        grad[v] = gx;
        grad[N + v] = gy;
        grad[2 * N + v] = gz;
    }
}

// VectorDomain::update_grad (src/base/vec_dom.rs:146-178); value 3N SoA -> grad 9N SoA
__global__ void __launch_bounds__(BLK) k_grad_vec(int64_t N, int64_t E, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ngh, const double* __restrict__ gw,
                                                  const double* __restrict__ val, double* __restrict__ grad) {
    int64_t v = (int64_t)xcd_block() * BLK + threadIdx.x;
    if (v >= N) return;
    double uc[3] = {val[v], val[N + v], val[2 * N + v]};
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t e = rp[v]; e < rp[v + 1]; ++e) {
        int32_t n = ngh[e];
        double wx = gw[e], wy = gw[E + e], wz = gw[2 * E + e];
        for (int c = 0; c < 3; ++c) {
            double d = val[c * N + n] - uc[c];
            g[3 * c] += wx * d;
            g[3 * c + 1] += wy * d;
            g[3 * c + 2] += wz * d;
        }
    }
    for (int k = 0; k < 9; ++k) grad[k * N + v] = g[k];
}

// flow_helper::update_flux face mass flux (src/operator/core/flow_helper.rs:118-171)
// for a prescribed velocity with uniform pressure and zero body force, where the
// Rhie-Chow velocity vel_rc vanishes (flow_helper.rs:165-167).
__global__ void __launch_bounds__(BLK) k_flux_prescribed(int64_t N, int64_t E, const int32_t* __restrict__ erow,
                                                         const int32_t* __restrict__ ngh,
                                                         const double* __restrict__ area, const double* __restrict__ norm,
                                                         const double* __restrict__ skew,
                                                         const double* __restrict__ den, const double* __restrict__ dgrad,
                                                         const double* __restrict__ vel, const double* __restrict__ vgrad,
                                                         double* __restrict__ flux) {
    int64_t e = (int64_t)xcd_block() * BLK + threadIdx.x;
    if (e >= E) return;
    const int64_t v = erow[e], n = ngh[e];
    const double sx = skew[e], sy = skew[E + e], sz = skew[2 * E + e];
    const double nx = norm[e], ny = norm[E + e], nz = norm[2 * E + e];
    double den_f = 0.5 * (den[v] + den[n]) +
                   0.5 * (sx * (dgrad[v] + dgrad[n]) + sy * (dgrad[N + v] + dgrad[N + n]) + sz * (dgrad[2 * N + v] + dgrad[2 * N + n]));
    double uf[3];
    for (int c = 0; c < 3; ++c) {
        const double* gc = vgrad + 3 * c * N;
        uf[c] = 0.5 * (vel[c * N + v] + vel[c * N + n]) +
                0.5 * (sx * (gc[v] + gc[n]) + sy * (gc[N + v] + gc[N + n]) + sz * (gc[2 * N + v] + gc[2 * N + n]));
    }
    const double vel_rc = 0.0;
    flux[e] = den_f * (uf[0] * nx + uf[1] * ny + uf[2] * nz + vel_rc) * area[e];
}

struct AsmArgs {
    int64_t N, E, S;
    const int32_t* rp;
    const int32_t* ngh;
    const double* dist;
    const double* area;
    const double* skew;  // 3E
    const double* vol;
    const double* cskew;  // 3N
    const int32_t* bsf_ptr;
    const int32_t* bsf_vj;
    const int32_t* bsf_vk;
    const double* bsf_area;
    const double* bsf_wgt;  // 3S
    const int32_t* vbnd;
    const int32_t* bidx;
    const int32_t* vcls;
    int32_t cls_itf0;  // first interface class (rows of k_assemble_itf)
    const int32_t* mrow;
    const int32_t* eslot;
    const int32_t* dslot;
    const DevProgram* prog;
    double dt;
    double* vals;
    double* b;
};

// harmonic mean, skew correction and positivity guard (opscl_dom_diff.rs:78-80)
__device__ inline double face_diff(const double* __restrict__ G, const double* __restrict__ Gg, int64_t N, int64_t v,
                                   int64_t n, double sx, double sy, double sz) {
    const double gv = G[v], gn = G[n];
    const double diff_m = (gv + gn != 0.0) ? 2.0 * gv * gn / (gv + gn) : 0.0;
    double diff_f = diff_m + 0.5 * (sx * (Gg[v] + Gg[n]) + sy * (Gg[N + v] + Gg[N + n]) + sz * (Gg[2 * N + v] + Gg[2 * N + n]));
    return diff_f > 0.0 ? diff_f : diff_m;
}

// Matrix coefficient of directed edge e = (v, n) for an edge operator.
__device__ inline double edge_coeff(const DevOp& op, const AsmArgs& a, int64_t v, int64_t e) {
    const int64_t N = a.N, E = a.E;
    const int64_t n = a.ngh[e];
    if (op.kind == VFVM_OP_SCL_DOM_DIFFUSION || op.kind == VFVM_OP_SCL_BND_DIFFUSION) {
        // opscl_dom_diff.rs:67-86, opscl_bnd_diff.rs:72-91
        double df = face_diff(op.f_val, op.f_grad, N, v, n, a.skew[e], a.skew[E + e], a.skew[2 * E + e]);
        return df * a.area[e] / a.dist[e];
    }
    // opscl_dom_advunity.rs:70-112: power law A = max(0, 1 - 0.1|Pe|)^5; opscl_dom_adv.rs:86-118
    // converts the mass flux with the skew-corrected face means of m and rho
    const double sx = a.skew[e], sy = a.skew[E + e], sz = a.skew[2 * E + e];
    double df = op.has_diff ? face_diff(op.f_val, op.f_grad, N, v, n, sx, sy, sz) : 0.0;
    double cond = df * a.area[e] / a.dist[e];
    double flux = op.flux[e];
    if (op.kind == VFVM_OP_SCL_DOM_ADVECTION || op.kind == VFVM_OP_SCL_BND_ADVECTION) {
        const double* W = op.w_val;
        const double* Wg = op.w_grad;
        const double* R = op.r_val;
        const double* Rg = op.r_grad;
        const double wgt_f = 0.5 * (W[v] + W[n]) + 0.5 * (sx * (Wg[v] + Wg[n]) + sy * (Wg[N + v] + Wg[N + n]) + sz * (Wg[2 * N + v] + Wg[2 * N + n]));
        const double den_f = 0.5 * (R[v] + R[n]) + 0.5 * (sx * (Rg[v] + Rg[n]) + sy * (Rg[N + v] + Rg[N + n]) + sz * (Rg[2 * N + v] + Rg[2 * N + n]));
        flux = den_f != 0.0 ? wgt_f / den_f * flux : 0.0;
    }
    double pe = cond != 0.0 ? flux / cond : 0.0;
    double base = fmax(1.0 - 0.1 * fabs(pe), 0.0);
    double b2 = base * base;
    double pl = b2 * b2 * base;
    return cond * (pl - 1.0) + fmax(-flux, 0.0);
}

__device__ inline bool is_edge_op(int k) {
    return k == VFVM_OP_SCL_DOM_DIFFUSION || k == VFVM_OP_SCL_BND_DIFFUSION || k == VFVM_OP_SCL_DOM_ADVECTION_UNITY ||
           k == VFVM_OP_SCL_BND_ADVECTION_UNITY || k == VFVM_OP_SCL_DOM_ADVECTION || k == VFVM_OP_SCL_BND_ADVECTION;
}

// Diagonal and right-hand-side contributions of a per-vertex operator.
__device__ inline void vertex_op(const DevOp& op, const AsmArgs& a, int64_t v, double& dacc, double& bacc) {
    const int64_t N = a.N, S = a.S;
    switch (op.kind) {
        case VFVM_OP_SCL_DOM_PRODUCTION:
        case VFVM_OP_SCL_BND_PRODUCTION: {
            // opscl_dom_prod.rs:56-72
            const double sx = a.cskew[v], sy = a.cskew[N + v], sz = a.cskew[2 * N + v];
            const double vol = a.vol[v];
            const double pv = op.f_val[v];
            const double prd_c = pv + op.f_grad[v] * sx + op.f_grad[N + v] * sy + op.f_grad[2 * N + v] * sz;
            const double unk_skew = op.u_grad[v] * sx + op.u_grad[N + v] * sy + op.u_grad[2 * N + v] * sz;
            dacc += prd_c * vol;
            bacc += -pv * unk_skew * vol;
        } break;
        case VFVM_OP_SCL_DOM_SOURCE:
        case VFVM_OP_SCL_BND_SOURCE: {
            // opscl_dom_src.rs:50-64
            const double sx = a.cskew[v], sy = a.cskew[N + v], sz = a.cskew[2 * N + v];
            const double src_c = op.f_val[v] + op.f_grad[v] * sx + op.f_grad[N + v] * sy + op.f_grad[2 * N + v] * sz;
            bacc += -src_c * a.vol[v];
        } break;
        case VFVM_OP_SCL_DOM_TIME:
        case VFVM_OP_SCL_BND_TIME: {
            // opscl_dom_time.rs:46-58
            const double vol = a.vol[v];
            const double w = op.f_val[v];
            dacc += -w * vol / a.dt;
            bacc += -w * op.u_prev[v] * vol / a.dt;
        } break;
        case VFVM_OP_SCL_DOM_TIME_PREVIOUS:
        case VFVM_OP_SCL_BND_TIME_PREVIOUS: {
            // opscl_dom_timeprev.rs:36-52: the weight's previous values
            const double vol = a.vol[v];
            const double w = op.f_val[v];
            dacc += -w * vol / a.dt;
            bacc += -w * op.u_prev[v] * vol / a.dt;
        } break;
        case VFVM_OP_SCL_DOM_TIME_UNITY:
        case VFVM_OP_SCL_BND_TIME_UNITY: {
            // opscl_dom_timeunity.rs:35-50
            const double vol = a.vol[v];
            dacc += -vol / a.dt;
            bacc += -op.u_prev[v] * vol / a.dt;
        } break;
        case VFVM_OP_SCL_BND_DIRICHLET: {
            // opscl_bnd_dir.rs:65-75
            dacc += 1.0;
            bacc += op.f_val[a.bidx[v]];
        } break;
        case VFVM_OP_SCL_BND_NEUMANN: {
            // opscl_bnd_neu.rs:54-83
            const double fv = op.f_val[a.bidx[v]];
            const int32_t vb = a.vbnd[v];
            for (int32_t s = a.bsf_ptr[v]; s < a.bsf_ptr[v + 1]; ++s) {
                const int32_t j = a.bsf_vj[s], k = a.bsf_vk[s];
                const double fj = a.vbnd[j] == vb ? op.f_val[a.bidx[j]] : fv;
                const double fk = a.vbnd[k] == vb ? op.f_val[a.bidx[k]] : fv;
                bacc += a.bsf_area[s] * (a.bsf_wgt[s] * fv + a.bsf_wgt[S + s] * fj + a.bsf_wgt[2 * S + s] * fk);
            }
        } break;
        default:
            break;
    }
}

// One row per LPR lanes: lane l handles the directed edges es + l, es + l + LPR (rows of
// degree <= 2 LPR), so the per-edge geometry and coefficient loads of a wave are
// contiguous.  Each matrix entry is summed in operator order in its lane's registers
// (starting from 0, as the reference's duplicate summation does); the diagonal is summed
// in the reference's push order (operator, then neighbour) by walking the lanes' edge
// coefficients with shuffles, so the result is bit-identical to a sequential row walk.
// Rows of higher degree take the sequential walk in their first lane.
// The row of vertex v for the operator class cls (shared by the wave-uniform and the
// per-lane paths of k_assemble).
template <int LPR>
__device__ inline void assemble_row(const AsmArgs& a, int64_t v, int lane, int cls) {
    const int nops = a.prog->nops[cls];
    const int32_t es = a.rp[v], ee = a.rp[v + 1];
    const int32_t deg = ee - es;
    double dacc = 0.0, bacc = 0.0;
    if (deg <= 2 * LPR) {
        const int32_t e0 = es + lane, e1 = es + LPR + lane;
        const bool h0 = e0 < ee, h1 = e1 < ee;
        double v0 = 0.0, v1 = 0.0;
        bool any_edge = false;
        for (int o = 0; o < nops; ++o) {
            const DevOp op = a.prog->ops[cls][o];
            if (is_edge_op(op.kind)) {
                any_edge = true;
                const double c0 = h0 ? edge_coeff(op, a, v, e0) : 0.0;
                const double c1 = h1 ? edge_coeff(op, a, v, e1) : 0.0;
                v0 += c0;
                v1 += c1;
                const int n0 = deg < LPR ? deg : LPR;
                for (int k = 0; k < n0; ++k) dacc += -__shfl(c0, k, LPR);
                for (int k = 0; k < deg - LPR; ++k) dacc += -__shfl(c1, k, LPR);
            } else if (lane == 0) {
                vertex_op(op, a, v, dacc, bacc);
            }
        }
        if (any_edge) {
            if (h0) a.vals[a.eslot[e0]] = v0;
            if (h1) a.vals[a.eslot[e1]] = v1;
        }
    } else if (lane == 0) {
        const int32_t ms = a.mrow[v], me = a.mrow[v + 1];
        for (int32_t k = ms; k < me; ++k) a.vals[k] = 0.0;
        for (int o = 0; o < nops; ++o) {
            const DevOp op = a.prog->ops[cls][o];
            if (is_edge_op(op.kind)) {
                for (int32_t e = es; e < ee; ++e) {
                    const double coeff = edge_coeff(op, a, v, e);
                    a.vals[a.eslot[e]] += coeff;
                    dacc += -coeff;
                }
            } else {
                vertex_op(op, a, v, dacc, bacc);
            }
        }
    }
    if (lane == 0) {
        const int32_t ds = a.dslot[v];
        if (ds >= 0) a.vals[ds] = dacc;
        a.b[v] = bacc;
    }
}

// One row per LPR lanes: lane l handles the directed edges es + l, es + l + LPR (rows of
// degree <= 2 LPR), so the per-edge geometry and coefficient loads of a wave are
// contiguous.  Each matrix entry is summed in operator order in its lane's registers
// (starting from 0, as the reference's duplicate summation does); the diagonal is summed
// in the reference's push order (operator, then neighbour) by walking the lanes' edge
// coefficients with shuffles, so the result is bit-identical to a sequential row walk.
// Rows of higher degree take the sequential walk in their first lane.  Rows are ordered
// points first, then the nodes of each boundary, so almost every wave holds one operator
// class: the class is then taken from the first lane (wave-uniform), and the operator
// descriptors are read once per wave through the scalar cache instead of once per lane.
template <int LPR>
__global__ void __launch_bounds__(BLK) k_assemble(AsmArgs a) {
    const int64_t tid = (int64_t)xcd_block() * BLK + threadIdx.x;
    const int64_t v = tid / LPR;
    const int lane = (int)(tid % LPR);
    if (v >= a.N) return;  // whole row groups leave together (BLK is a multiple of LPR)
    const int cls = a.vcls[v];
    if (cls >= a.cls_itf0) return;  // interface rows: k_assemble_itf
    const int cls0 = __builtin_amdgcn_readfirstlane(cls);
    if (__all(cls == cls0)) assemble_row<LPR>(a, v, lane, cls0);
    else assemble_row<LPR>(a, v, lane, cls);
}

// The rows of the interface nodes' copies (joined meshes), one thread per node q: every push
// of the interface operators (opscl_itf_*.rs) is added into its slot in the reference's push
// order (operator, then side a before side b, then face order), so each entry and b are the
// reference's sequential sums.  Side a's row takes the continuity operator, side b's row every
// other one; a face coefficient is the domain kernel's (edge_coeff) on the face of that side's
// copy of the node.
struct ItfArgs {
    int64_t K;
    const int32_t* va;
    const int32_t* vb;
    const int32_t* itf;
    const int32_t* islot;  // 4 per node: (a,a), (a,b) of row a_v; (b,a), (b,b) of row b_v
};
__device__ inline int itf_face_kind(int k) {
    return k == VFVM_OP_SCL_ITF_DIFFUSION ? VFVM_OP_SCL_DOM_DIFFUSION
         : k == VFVM_OP_SCL_ITF_ADVECTION ? VFVM_OP_SCL_DOM_ADVECTION
                                           : VFVM_OP_SCL_DOM_ADVECTION_UNITY;
}
__global__ void __launch_bounds__(BLK) k_assemble_itf(AsmArgs a, ItfArgs f) {
    const int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (q >= f.K) return;
    const int64_t N = a.N;
    const int32_t va = f.va[q], vb = f.vb[q];
    const int clsA = a.cls_itf0 + 2 * f.itf[q], clsB = clsA + 1;
    const int32_t saa = f.islot[4 * q], sab = f.islot[4 * q + 1], sba = f.islot[4 * q + 2], sbb = f.islot[4 * q + 3];
    // side a's row (opscl_itf_cont.rs:44-56): c_a - c_b = 0
    for (int32_t k = a.mrow[va]; k < a.mrow[va + 1]; ++k) a.vals[k] = 0.0;
    for (int o = 0; o < a.prog->nops[clsA]; ++o) {
        a.vals[saa] += 1.0;
        a.vals[sab] += -1.0;
    }
    a.b[va] = 0.0;
    // side b's row: the combined control volume
    for (int32_t k = a.mrow[vb]; k < a.mrow[vb + 1]; ++k) a.vals[k] = 0.0;
    double bacc = 0.0;
    for (int o = 0; o < a.prog->nops[clsB]; ++o) {
        DevOp op = a.prog->ops[clsB][o];
        const int kind = op.kind;
        for (int side = 0; side < 2; ++side) {
            const int32_t sv = side ? vb : va;
            const int32_t sdiag = side ? sbb : sba;
            switch (kind) {
                case VFVM_OP_SCL_ITF_DIFFUSION:
                case VFVM_OP_SCL_ITF_ADVECTION:
                case VFVM_OP_SCL_ITF_ADVECTION_UNITY: {
                    // opscl_itf_diff.rs:73-108 (both sides); opscl_itf_adv.rs:86-148 and
                    // opscl_itf_advunity.rs:85-135 (the flowing side only)
                    if (kind != VFVM_OP_SCL_ITF_DIFFUSION && side != op.side) break;
                    op.kind = itf_face_kind(kind);
                    for (int32_t e = a.rp[sv]; e < a.rp[sv + 1]; ++e) {
                        const double coeff = edge_coeff(op, a, sv, e);
                        a.vals[a.eslot[e]] += coeff;
                        a.vals[sdiag] += -coeff;
                    }
                    op.kind = kind;
                } break;
                case VFVM_OP_SCL_ITF_PRODUCTION: {
                    // opscl_itf_prod.rs:60-92: both halves on (b_v, b_v)
                    const double sx = a.cskew[sv], sy = a.cskew[N + sv], sz = a.cskew[2 * N + sv];
                    const double vol = a.vol[sv];
                    const double pv = op.f_val[sv];
                    const double prd_c = pv + op.f_grad[sv] * sx + op.f_grad[N + sv] * sy + op.f_grad[2 * N + sv] * sz;
                    const double unk_skew = op.u_grad[sv] * sx + op.u_grad[N + sv] * sy + op.u_grad[2 * N + sv] * sz;
                    a.vals[sbb] += prd_c * vol;
                    bacc += -pv * unk_skew * vol;
                } break;
                case VFVM_OP_SCL_ITF_SOURCE: {
                    // opscl_itf_src.rs:57-83
                    const double sx = a.cskew[sv], sy = a.cskew[N + sv], sz = a.cskew[2 * N + sv];
                    const double src_c = op.f_val[sv] + op.f_grad[sv] * sx + op.f_grad[N + sv] * sy + op.f_grad[2 * N + sv] * sz;
                    bacc += -src_c * a.vol[sv];
                } break;
                case VFVM_OP_SCL_ITF_TIME:
                case VFVM_OP_SCL_ITF_TIME_PREVIOUS: {
                    // opscl_itf_time.rs:50-76, opscl_itf_timeprev.rs:50-78: each side's term on
                    // that side's copy of the unknown
                    const double vol = a.vol[sv];
                    const double w = op.f_val[sv];
                    a.vals[sdiag] += -w * vol / a.dt;
                    bacc += -w * op.u_prev[sv] * vol / a.dt;
                } break;
                case VFVM_OP_SCL_ITF_TIME_UNITY: {
                    // opscl_itf_timeunity.rs:47-70
                    const double vol = a.vol[sv];
                    a.vals[sdiag] += -vol / a.dt;
                    bacc += -op.u_prev[sv] * vol / a.dt;
                } break;
                default:
                    break;
            }
        }
    }
    a.b[vb] = bacc;
}

}  // namespace

int launch_grad(Ctx& c, const double* value, double* grad) {
    DevGeom& g = c.g;
    if (g.N == 0) return VFVM_OK;
    constexpr int LPR = 16;
    hipLaunchKernelGGL((k_grad<LPR>), dim3(grid_for(g.N * LPR)), dim3(BLK), 0, c.stream, g.N, g.E, g.row_ptr.p, g.ngh.p, g.gw.p, value, grad);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

int launch_grad_vec(Ctx& c, const double* value3, double* grad9) {
    DevGeom& g = c.g;
    if (g.N == 0) return VFVM_OK;
    hipLaunchKernelGGL(k_grad_vec, dim3(grid_for(g.N)), dim3(BLK), 0, c.stream, g.N, g.E, g.row_ptr.p, g.ngh.p, g.gw.p, value3, grad9);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

int launch_flux_prescribed(Ctx& c, SclFlw& f) {
    DevGeom& g = c.g;
    VFVM_HIP_TRY(f.flux.alloc(g.E));
    SclDom& den = c.scl[f.den];
    VecDom& vel = c.vec[f.vel];
    hipLaunchKernelGGL(k_flux_prescribed, dim3(grid_for(g.E)), dim3(BLK), 0, c.stream, g.N, g.E, g.erow.p, g.ngh.p,
                       g.area.p, g.norm.p, g.skew.p, den.value.p, den.grad.p, vel.value.p, vel.grad.p, f.flux.p);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

int launch_assemble(Ctx& c, double dt) {
    return launch_assemble_block(c, dt, c.mrow.p, c.eslot.p, c.dslot.p, c.prog.p, c.b.p);
}

int launch_assemble_block(Ctx& c, double dt, const int32_t* mrow, const int32_t* eslot, const int32_t* dslot,
                          const DevProgram* prog, double* b) {
    DevGeom& g = c.g;
    AsmArgs a;
    a.N = g.N; a.E = g.E; a.S = g.S;
    a.rp = g.row_ptr.p; a.ngh = g.ngh.p; a.dist = g.dist.p; a.area = g.area.p; a.skew = g.skew.p;
    a.vol = g.vol.p; a.cskew = g.cskew.p;
    a.bsf_ptr = g.bsf_ptr.p; a.bsf_vj = g.bsf_vj.p; a.bsf_vk = g.bsf_vk.p; a.bsf_area = g.bsf_area.p; a.bsf_wgt = g.bsf_wgt.p;
    a.vbnd = g.vbnd.p; a.bidx = g.bidx.p;
    a.vcls = g.vcls.p; a.cls_itf0 = g.cls_itf0;
    a.mrow = mrow; a.eslot = eslot; a.dslot = dslot;
    a.prog = prog;
    a.dt = dt;
    a.vals = c.vals.p; a.b = b;
    // algorithmic bytes: per directed edge ngh, eslot (4+4), area, dist, skew (5 x 8), one value
    // write (8) and the neighbour's coefficient value + gradient (32); per vertex CSR offsets,
    // classes, geometry, b (~64)
    double bytes = (double)g.E * (8.0 + 40.0 + 8.0 + 32.0) + (double)g.N * 64.0;
    hipEvent_t ev = nullptr;
    ktime_begin(c, VFVM_K_ASSEMBLE, &ev);
    constexpr int LPR = 16;
    hipLaunchKernelGGL((k_assemble<LPR>), dim3(grid_for(g.N * LPR)), dim3(BLK), 0, c.stream, a);
    VFVM_HIP_TRY(hipGetLastError());
    if (g.K > 0 && mrow == c.mrow.p) {
        ItfArgs f{g.K, g.itf_va.p, g.itf_vb.p, g.itf_of.p, c.islot.p};
        hipLaunchKernelGGL(k_assemble_itf, dim3(grid_for(g.K)), dim3(BLK), 0, c.stream, a, f);
        VFVM_HIP_TRY(hipGetLastError());
    }
    ktime_end(c, VFVM_K_ASSEMBLE, bytes, ev);
    return VFVM_OK;
}

}  // namespace vfvm
