// Discrete adjoint of the Picard-converged coupled flow system (no reference counterpart).
//
// A state x = [u, v, w, p, s_1, s_2] of a flow list (SteadyFlow / TransientFlow /
// TransientHeat(Mass)Flow) is a fixed point of the Picard map: R(x) = A(c(x)) x - b(c(x)) = 0,
// where c(x) are the state-dependent coefficients the flow update computes from x
// (flow_helper.rs:3-179): the face mass flux F (Rhie-Chow), rc = V / a_P, the velocity
// gradients, and the body force (an affine device closure fce = f0 + df s).  The Jacobian is
// J = A + E, E = (dR/dc)(dc/dx) at fixed x.  J^T l = A^T l (the transposed CSR) + E^T l, and
// E^T l is a chain of vector-Jacobian products evaluated by three gather kernels:
//
//   edges   (k_adj_edge):   gF_e = d(l . R)/dF_e through the momentum advection rows
//                           (opvec_dom_adv.rs:72-108) and the scalar advection rows
//                           (opscl_dom_adv.rs:86-118, opscl_dom_advunity.rs:70-112):
//                           l_v phi'(F_e) (x_n - x_v) per row, phi the power-law coefficient;
//                           wF_e = gF_e rho_f A_e (adjoint of the face velocity sum
//                           uf . n + vel_rc) and wC_e = -l_p,v rho_f A_e: the continuity row of
//                           v is -sum_e F_e (divergence + pressure stabiliser,
//                           opscl_dom_div.rs:86-150, opscl_dom_pstab.rs:70-148), whose
//                           dependence on rc and the body force E carries (its pressure and
//                           velocity columns are in A).
//   vertices (k_adj_vertex): the adjoint of vel_rc_e = -(rc_v + rc_n)/2 ((p_n - p_v)/d - fce_e)
//                           + n . (mom_v + mom_n)/2, mom = rc (grad p - fce_c), and of
//                           uf = (u_v + u_n)/2 + skew . (grad u_v + grad u_n)/2 over each
//                           vertex's own and reverse edges (flow_helper.rs:114-171); the
//                           transpose / dilatatory / source right-hand sides of the momentum
//                           rows (opvec_dom_dtrn.rs:70-104, opvec_dom_ddil.rs:74-103,
//                           opvec_dom_src.rs:51-70); d rc / d a_P = -rc^2 / V.
//   final   (k_adj_final):  a_P = sum_e cond P(|Fc_e / cond|) + max(Fc_e, 0) with the central
//                           flux Fc (flow_helper.rs:3-70); the least-squares gradients
//                           (vec_dom.rs:146-178) of the velocity and the body force
//                           transposed through the reverse edge map; grad p =
//                           sum_q gc_q p_q / V transposed through the coefficient lists'
//                           transpose; the body force's closure df into its scalar unknown.
//
// Each output entry is a fixed-order sum in one thread: no atomics, run-to-run identical.
// The backward pass solves J^T l^n = w_n seed - D l^(n+1) (D = dR^(n+1)/dx^n: rho V / dt on
// the momentum rows, w V / dt on the scalar rows with time operators) per state by GCR with
// the transposed GMRES on A^T as its preconditioner, replaying each checkpoint segment with
// the forward driver to recover the states (bit-identical: the path has no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "flow.hpp"

using namespace vfvm;

namespace {

constexpr int BLK = 256;
constexpr int DOT_BLOCKS = 256;
constexpr double NEG_FRAC_TWO_THREE = -2.0 / 3.0;  // opvec_dom_ddil.rs:6
inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

__device__ inline bool is_adv(int k) {
    return k == VFVM_OP_SCL_DOM_ADVECTION || k == VFVM_OP_SCL_BND_ADVECTION || k == VFVM_OP_SCL_DOM_ADVECTION_UNITY ||
           k == VFVM_OP_SCL_BND_ADVECTION_UNITY;
}

// d/dF of the advection coefficient phi(F) = cond (P(F / cond) - 1) + max(-F, 0),
// P(x) = max(0, 1 - 0.1 |x|)^5: P'(Pe) - H(-F) = -0.5 a^4 sign(Pe) - H(-F) with
// a = max(0, 1 - 0.1 |Pe|), Pe = F / cond.  The sign is that of Pe, not of F: a dual face
// next to the boundary can have a negative area, and then cond < 0.  At F = 0 the value is
// -0.5, the derivative for cond > 0 and the mean of the one-sided ones for cond < 0.  With
// cond = 0 the coefficient is max(-F, 0).
__device__ inline double dphi_adv(double F, double cond) {
    if (cond == 0.0) return F < 0.0 ? -1.0 : 0.0;
    if (F == 0.0) return -0.5;
    const double pe = F / cond;
    const double a = fmax(1.0 - 0.1 * fabs(pe), 0.0);
    const double a4 = (a * a) * (a * a);
    return (pe > 0.0 ? -0.5 * a4 : 0.5 * a4) - (F < 0.0 ? 1.0 : 0.0);
}
// d/dF of h(F) = cond P(F / cond) + max(F, 0), the face term of a_P: P'(Pe) + H(F); 0.5 at
// F = 0 (as above)
__device__ inline double dh_ap(double F, double cond) {
    if (cond == 0.0) return F > 0.0 ? 1.0 : 0.0;
    if (F == 0.0) return 0.5;
    const double pe = F / cond;
    const double a = fmax(1.0 - 0.1 * fabs(pe), 0.0);
    const double a4 = (a * a) * (a * a);
    return (pe > 0.0 ? -0.5 * a4 : 0.5 * a4) + (F > 0.0 ? 1.0 : 0.0);
}

struct AdjBlk {
    const DevProgram* prog;  // the scalar unknown's operators per vertex class
    const double* s;         // its values
};

// The adjoint vectors (l, y, the outer solve's directions) and the transposed system are
// numbered vertex-interleaved: unknown c of vertex v at S v + c, S = 4 + coupled scalars, so
// that a row's gathers of a neighbour's unknowns share cache lines.  The work arrays of one
// product are vertex-major (the 3 or 9 entries of a vertex adjacent).
struct FlowAdjArgs {
    FlowGeom g;
    FlowFields f;
    int64_t pref_v;  // vertex of the reference-pressure row (-1: none)
    int32_t S;       // unknowns per vertex
    int32_t nblk;
    int32_t body_blk;  // block whose unknown drives the body force (-1: the force is fixed)
    AdjBlk blk[2];
    const double* dfce;  // 3N: d fce / d s of the affine closure
    const int32_t* rev;
    const int32_t* gct_ptr;
    const int32_t* gct_slot;
    const int32_t* gct_own;
    // state-only factors (k_adj_coeffs): per directed edge, 7E SoA
    double* ec;
    // per vertex: pg = grad p - fce_c (3N, vertex-major), sgw = sum of the edges' gradient
    // weights (3N, vertex-major)
    double* pg;
    double* sgw;
    const double* lam;  // adjoint vector, entries (u v w p s1 s2) of each vertex adjacent
    double* wF;         // E
    double* wC;         // E
    double* gu;         // 3N
    double* gp1;        // N
    double* ggp;        // 3N: adjoint of grad p
    double* gaP;        // N: adjoint of a_P
    double* gvelg;      // 9N
    double* gfce;       // 3N
    double* gfceg;      // 9N
    const double* base; // A^T l (added to the output)
    double* y;          // n
};

// rows of the per-edge factors in ec (entry k E + e)
enum { EC_RHOA = 0, EC_MPHI = 1, EC_SPHI0 = 2, EC_SPHI1 = 3, EC_MU2A = 4, EC_APW = 5, EC_TFAC = 6, EC_ROWS = 7 };

__device__ inline int64_t ix(const FlowAdjArgs& p, int c, int64_t v) { return (int64_t)p.S * v + c; }

constexpr int LPV = 16;  // lanes per vertex of the gathers

// sum over the LPV lanes of a vertex, in a fixed tree order (the total in every lane)
__device__ inline double lane_sum(double x) {
#pragma unroll
    for (int off = LPV / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, LPV);
    return x;
}

// Per directed edge e = (v, n), once per state: rho_f A; the momentum advection's slope
// d phi / dF (rows of the points); per coupled scalar the slope sum of its advection operators
// kappa phi'(kappa F); mu_f A / 2 of the transpose / dilatatory sums (points); the a_P factor
// h'(Fc) rho_f A / 2; the Rhie-Chow term -((p_n - p_v) / d - fce_e) / 2.
__global__ void __launch_bounds__(BLK) k_adj_coeffs(FlowAdjArgs p) {
    const int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
    const FlowGeom& g = p.g;
    const FlowFields& f = p.f;
    if (e >= g.E) return;
    const int64_t N = g.N, E = g.E;
    const int64_t v = g.erow[e], n = g.ngh[e];
    const double A = g.area[e], d = g.dist[e];
    const double nx = g.norm[e], ny = g.norm[E + e], nz = g.norm[2 * E + e];
    const double sx = g.skew[e], sy = g.skew[E + e], sz = g.skew[2 * E + e];
    const double den_f = face_mean(f.den, f.deng, N, v, n, sx, sy, sz);
    const double visc_f = face_visc(f.visc, f.viscg, N, v, n, sx, sy, sz);
    const double cond = visc_f * A / d;
    const double F = f.flux[e];
    const int32_t vb = g.vbnd[v];
    const int cls = vb < 0 ? 0 : 1 + vb;
    double sphi[2] = {0.0, 0.0};
    for (int k = 0; k < p.nblk; ++k) {
        const DevProgram* prog = p.blk[k].prog;
        for (int o = 0; o < prog->nops[cls]; ++o) {
            const DevOp& op = prog->ops[cls][o];
            if (!is_adv(op.kind)) continue;
            const double df = op.has_diff ? face_visc(op.f_val, op.f_grad, N, v, n, sx, sy, sz) : 0.0;
            double kap = 1.0;
            if (op.kind == VFVM_OP_SCL_DOM_ADVECTION || op.kind == VFVM_OP_SCL_BND_ADVECTION) {
                const double wgt_f = face_mean(op.w_val, op.w_grad, N, v, n, sx, sy, sz);
                const double rho_f = face_mean(op.r_val, op.r_grad, N, v, n, sx, sy, sz);
                kap = rho_f != 0.0 ? wgt_f / rho_f : 0.0;
            }
            sphi[k] += kap * dphi_adv(kap * op.flux[e], df * A / d);
        }
    }
    const double Fc = den_f * 0.5 * ((f.vel[v] + f.vel[n]) * nx + (f.vel[N + v] + f.vel[N + n]) * ny + (f.vel[2 * N + v] + f.vel[2 * N + n]) * nz) * A;
    const double fce_e = 0.5 * (nx * (f.fce[v] + f.fce[n]) + ny * (f.fce[N + v] + f.fce[N + n]) + nz * (f.fce[2 * N + v] + f.fce[2 * N + n]));
    double* ec = p.ec;
    ec[EC_RHOA * E + e] = den_f * A;
    ec[EC_MPHI * E + e] = v < g.NP ? dphi_adv(F, cond) : 0.0;
    ec[EC_SPHI0 * E + e] = sphi[0];
    ec[EC_SPHI1 * E + e] = sphi[1];
    ec[EC_MU2A * E + e] = v < g.NP ? 0.5 * visc_f * A : 0.0;
    ec[EC_APW * E + e] = dh_ap(Fc, cond) * den_f * 0.5 * A;
    ec[EC_TFAC * E + e] = -0.5 * ((f.pres[n] - f.pres[v]) / d - fce_e);
}

// Per vertex, once per state: pg_v = grad p_v - fce_c,v with grad p_v = sum_q gc_q p_q / V_v
// (the factors of mom_v / rc_v, flow_helper.rs:85-112), and the sums of the edges' gradient
// weights (the vertex's own term of the transposed least-squares gradient)
__global__ void __launch_bounds__(BLK) k_adj_vcoeffs(FlowAdjArgs p) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    const FlowGeom& g = p.g;
    if (v >= g.N) return;
    const int64_t E = g.E, G = g.G;
    const double vol = g.vol[v];
    double px = 0.0, py = 0.0, pz = 0.0;
    for (int32_t q = g.gc_ptr[v]; q < g.gc_ptr[v + 1]; ++q) {
        const double pr = p.f.pres[g.gc_vid[q]];
        px += g.gc_c[q] * pr / vol;
        py += g.gc_c[G + q] * pr / vol;
        pz += g.gc_c[2 * G + q] * pr / vol;
    }
    double fc[3];
    fce_centroid(g, p.f, v, fc);
    p.pg[3 * v] = px - fc[0];
    p.pg[3 * v + 1] = py - fc[1];
    p.pg[3 * v + 2] = pz - fc[2];
    double wx = 0.0, wy = 0.0, wz = 0.0;
    for (int32_t e = g.rp[v]; e < g.rp[v + 1]; ++e) {
        wx += g.gw[e];
        wy += g.gw[E + e];
        wz += g.gw[2 * E + e];
    }
    p.sgw[3 * v] = wx;
    p.sgw[3 * v + 1] = wy;
    p.sgw[3 * v + 2] = wz;
}

// per directed edge e = (v, n): wF_e, wC_e
__global__ void __launch_bounds__(BLK) k_adj_edge(FlowAdjArgs p) {
    const int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
    const FlowGeom& g = p.g;
    const FlowFields& f = p.f;
    if (e >= g.E) return;
    const int64_t N = g.N, E = g.E;
    const int64_t v = g.erow[e], n = g.ngh[e];
    const double* ec = p.ec;
    double gF = 0.0;
    const double mphi = ec[EC_MPHI * E + e];
    if (mphi != 0.0) {
        double t = 0.0;
        for (int c = 0; c < 3; ++c) t += p.lam[ix(p, c, v)] * (f.vel[c * N + n] - f.vel[c * N + v]);
        gF += mphi * t;
    }
    for (int k = 0; k < p.nblk; ++k) {
        const double sphi = ec[(EC_SPHI0 + k) * E + e];
        if (sphi != 0.0) gF += sphi * p.lam[ix(p, 4 + k, v)] * (p.blk[k].s[n] - p.blk[k].s[v]);
    }
    const double rhoA = ec[EC_RHOA * E + e];
    p.wF[e] = gF * rhoA;
    p.wC[e] = (v == p.pref_v) ? 0.0 : -p.lam[ix(p, 3, v)] * rhoA;
}

// Per vertex u (LPV lanes, each over a stride of u's edges e = (u, n) and their reverses
// r = (n, u)): the adjoint of the face velocity uf . n and its skew term, of vel_rc (the
// pressure difference and grad p through mom_u on the flux path; rc_u and the body force on
// both paths) and of the momentum rows' transpose / dilatatory / source right-hand sides.
__global__ void __launch_bounds__(BLK) k_adj_vertex(FlowAdjArgs p) {
    const int64_t tid = (int64_t)blockIdx.x * BLK + threadIdx.x;
    const int64_t u = tid / LPV;
    const int lane = (int)(tid % LPV);
    const FlowGeom& g = p.g;
    const FlowFields& f = p.f;
    if (u >= g.N) return;  // whole vertex groups leave together (BLK is a multiple of LPV)
    const int64_t E = g.E;
    const double* ec = p.ec;
    const double rc_u = f.rc[u];
    const double pgu[3] = {p.pg[3 * u], p.pg[3 * u + 1], p.pg[3 * u + 2]};
    double gu[3] = {0.0, 0.0, 0.0}, ggp[3] = {0.0, 0.0, 0.0}, gfn[3] = {0.0, 0.0, 0.0}, wn[3] = {0.0, 0.0, 0.0};
    double gvelg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double gp = 0.0, grc = 0.0, gdiv = 0.0;
    for (int32_t e = g.rp[u] + lane; e < g.rp[u + 1]; e += LPV) {
        const int64_t n = g.ngh[e];
        const int64_t r = p.rev[e];
        const double rc_n = f.rc[n];
        for (int side = 0; side < 2; ++side) {
            // side 0: edge e = (u, n), u its first vertex; side 1: r = (n, u), u its second
            const int64_t k = side == 0 ? e : r;
            const double d = g.dist[k];
            const double nv[3] = {g.norm[k], g.norm[E + k], g.norm[2 * E + k]};
            const double sk[3] = {g.skew[k], g.skew[E + k], g.skew[2 * E + k]};
            const double wf = p.wF[k], wt = wf + p.wC[k];
            for (int c = 0; c < 3; ++c) {
                const double t = wf * 0.5 * nv[c];
                gu[c] += t;
                for (int q = 0; q < 3; ++q) gvelg[3 * c + q] += t * sk[q];
                ggp[c] += wf * 0.5 * rc_u * nv[c];
                gfn[c] += wt * (0.25 * (rc_u + rc_n) - 0.5 * rc_u) * nv[c];
                wn[c] += wt * nv[c];
            }
            gp += wf * (side == 0 ? 0.5 : -0.5) * (rc_u + rc_n) / d;
            grc += wt * (ec[EC_TFAC * E + k] + 0.5 * (nv[0] * pgu[0] + nv[1] * pgu[1] + nv[2] * pgu[2]));
            // transpose and dilatatory sums of the momentum row of the edge's first vertex (a
            // point) read the face mean of u's velocity gradient
            const double mu2A = ec[EC_MU2A * E + k];
            if (mu2A != 0.0) {
                const int64_t row_v = side == 0 ? u : n;
                for (int c = 0; c < 3; ++c) {
                    const double lc = p.lam[ix(p, c, row_v)];
                    for (int q = 0; q < 3; ++q) gvelg[3 * q + c] += lc * mu2A * nv[q];
                    gdiv += lc * (NEG_FRAC_TWO_THREE * mu2A * nv[c]);
                }
            }
        }
    }
    for (int c = 0; c < 3; ++c) {
        gu[c] = lane_sum(gu[c]);
        ggp[c] = lane_sum(ggp[c]);
        gfn[c] = lane_sum(gfn[c]);
        wn[c] = lane_sum(wn[c]);
    }
    for (int q = 0; q < 9; ++q) gvelg[q] = lane_sum(gvelg[q]);
    gp = lane_sum(gp);
    grc = lane_sum(grc);
    gdiv = lane_sum(gdiv);
    if (lane != 0) return;
    const double vol_u = g.vol[u];
    const double cs[3] = {g.cskew[u], g.cskew[g.N + u], g.cskew[2 * g.N + u]};
    // body force: vel_rc reads fce_e and the centroid values fc = fce + grad fce . cskew
    // (-rc_u / 2 n_c cskew_q per face); the point rows' source b -= fc V
    const double lv = u < g.NP ? 1.0 : 0.0;
    for (int c = 0; c < 3; ++c) {
        const double lc = lv * p.lam[ix(p, c, u)];
        p.gfce[3 * u + c] = gfn[c] + lc * vol_u;
        for (int q = 0; q < 3; ++q) p.gfceg[9 * u + 3 * c + q] = -0.5 * rc_u * wn[c] * cs[q] + lc * vol_u * cs[q];
        p.gu[3 * u + c] = gu[c];
        p.ggp[3 * u + c] = ggp[c];
    }
    gvelg[0] += gdiv;
    gvelg[4] += gdiv;
    gvelg[8] += gdiv;
    for (int q = 0; q < 9; ++q) p.gvelg[9 * u + q] = gvelg[q];
    p.gp1[u] = gp;
    p.gaP[u] = rc_u > 0.0 ? -grc * rc_u * rc_u / vol_u : 0.0;
}

// Per vertex u (LPV lanes): the a_P chain dFc/du through h'(Fc) (of u's own a_P over e and of
// each neighbour's over r), the transposed least-squares gradients of the velocity and the
// body force (grad_q x_c,m = sum over m's edges gw_q (x_c,o - x_c,m)), the transposed
// pressure gradient (grad p_v = sum_q gc_q p_q / V_v, over the coefficient lists' transpose),
// the body force's closure; y = A^T l + E^T l.
__global__ void __launch_bounds__(BLK) k_adj_final(FlowAdjArgs p) {
    const int64_t tid = (int64_t)blockIdx.x * BLK + threadIdx.x;
    const int64_t u = tid / LPV;
    const int lane = (int)(tid % LPV);
    const FlowGeom& g = p.g;
    if (u >= g.N) return;
    const int64_t N = g.N, E = g.E, G = g.G;
    const double* ec = p.ec;
    const double gaP_u = p.gaP[u];
    double gu[3] = {0.0, 0.0, 0.0}, gf[3] = {0.0, 0.0, 0.0}, gpv = 0.0;
    for (int32_t e = g.rp[u] + lane; e < g.rp[u + 1]; e += LPV) {
        const int64_t n = g.ngh[e];
        const int64_t r = p.rev[e];
        const double w0 = gaP_u * ec[EC_APW * E + e], w1 = p.gaP[n] * ec[EC_APW * E + r];
        const double wr[3] = {g.gw[r], g.gw[E + r], g.gw[2 * E + r]};
        for (int c = 0; c < 3; ++c) {
            gu[c] += w0 * g.norm[(int64_t)c * E + e] + w1 * g.norm[(int64_t)c * E + r];
            for (int q = 0; q < 3; ++q) {
                gu[c] += wr[q] * p.gvelg[9 * n + 3 * c + q];
                gf[c] += wr[q] * p.gfceg[9 * n + 3 * c + q];
            }
        }
    }
    for (int32_t t = p.gct_ptr[u] + lane; t < p.gct_ptr[u + 1]; t += LPV) {
        const int32_t q = p.gct_slot[t];
        const int64_t v = p.gct_own[t];
        gpv += (p.ggp[3 * v] * g.gc_c[q] + p.ggp[3 * v + 1] * g.gc_c[G + q] + p.ggp[3 * v + 2] * g.gc_c[2 * G + q]) / g.vol[v];
    }
    for (int c = 0; c < 3; ++c) {
        gu[c] = lane_sum(gu[c]);
        gf[c] = lane_sum(gf[c]);
    }
    gpv = lane_sum(gpv);
    if (lane != 0) return;
    const double sw[3] = {p.sgw[3 * u], p.sgw[3 * u + 1], p.sgw[3 * u + 2]};
    for (int c = 0; c < 3; ++c) {
        gu[c] += p.gu[3 * u + c];
        gf[c] += p.gfce[3 * u + c];
        for (int q = 0; q < 3; ++q) {
            gu[c] -= sw[q] * p.gvelg[9 * u + 3 * c + q];
            gf[c] -= sw[q] * p.gfceg[9 * u + 3 * c + q];
        }
    }
    for (int c = 0; c < 3; ++c) p.y[ix(p, c, u)] = p.base[ix(p, c, u)] + gu[c];
    p.y[ix(p, 3, u)] = p.base[ix(p, 3, u)] + gpv + p.gp1[u];
    for (int k = 0; k < p.nblk; ++k) {
        double s = 0.0;
        if (k == p.body_blk) s = p.dfce[u] * gf[0] + p.dfce[N + u] * gf[1] + p.dfce[2 * N + u] * gf[2];
        p.y[ix(p, 4 + k, u)] = p.base[ix(p, 4 + k, u)] + s;
    }
}

// D of dR^(n+1)/dx^n = +D: rho V / dt on the momentum rows of the points (the velocity time
// operator, opvec_dom_time.rs:51-56), the time operators' w V / dt on the scalar rows
// (opscl_dom_time.rs:46-58, opscl_dom_timeprev.rs:36-52, opscl_dom_timeunity.rs:35-50)
__global__ void __launch_bounds__(BLK) k_adj_time_diag(FlowAdjArgs p, double dt, double* __restrict__ D) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    const FlowGeom& g = p.g;
    if (v >= g.N) return;
    const double vol = g.vol[v];
    const double dm = v < g.NP ? p.f.den[v] * vol / dt : 0.0;
    for (int c = 0; c < 3; ++c) D[ix(p, c, v)] = dm;
    D[ix(p, 3, v)] = 0.0;
    const int32_t vb = g.vbnd[v];
    const int cls = vb < 0 ? 0 : 1 + vb;
    for (int k = 0; k < p.nblk; ++k) {
        const DevProgram* prog = p.blk[k].prog;
        double acc = 0.0;
        for (int o = 0; o < prog->nops[cls]; ++o) {
            const DevOp& op = prog->ops[cls][o];
            if (op.kind == VFVM_OP_SCL_DOM_TIME || op.kind == VFVM_OP_SCL_BND_TIME || op.kind == VFVM_OP_SCL_DOM_TIME_PREVIOUS ||
                op.kind == VFVM_OP_SCL_BND_TIME_PREVIOUS)
                acc += op.f_val[v] * vol / dt;
            else if (op.kind == VFVM_OP_SCL_DOM_TIME_UNITY || op.kind == VFVM_OP_SCL_BND_TIME_UNITY)
                acc += vol / dt;
        }
        D[ix(p, 4 + k, v)] = acc;
    }
}

// per-block partial sums of a . b (grid-stride, fixed order)
__global__ void __launch_bounds__(BLK) k_dot_part(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                  double* __restrict__ part) {
    __shared__ double s[BLK];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) acc += a[i] * b[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = BLK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
// per-block partial sums of x over the pressure entries but the reference vertex's
__global__ void __launch_bounds__(BLK) k_psum_part(int64_t N, int S, int64_t pref_v, const double* __restrict__ x,
                                                   double* __restrict__ part) {
    __shared__ double s[BLK];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLK)
        if (i != pref_v) acc += x[(int64_t)S * i + 3];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = BLK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
// y = indicator of the continuity rows but the reference one
__global__ void __launch_bounds__(BLK) k_cont_indicator(int64_t n, int S, int64_t pref_v, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) y[i] = (i % S == 3 && i / S != pref_v) ? 1.0 : 0.0;
}
// W -= t on the velocity entries of the nodes (their rows are identity rows of A)
__global__ void __launch_bounds__(BLK) k_mass_correct(int64_t N, int64_t NP, int S, const double* __restrict__ t,
                                                      double* __restrict__ W) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= 3 * (N - NP)) return;
    const int64_t k = (int64_t)S * (NP + i / 3) + i % 3;
    W[k] -= t[k];
}
// comp-major (entry (c, v) at c N + v) to vertex-interleaved (at S v + c), and back
__global__ void __launch_bounds__(BLK) k_to_il(int64_t N, int S, const double* __restrict__ in, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < (int64_t)S * N) out[(int64_t)S * (i % N) + i / N] = in[i];
}
__global__ void __launch_bounds__(BLK) k_from_il(int64_t N, int S, const double* __restrict__ in, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < (int64_t)S * N) out[i] = in[(int64_t)S * (i % N) + i / N];
}
// y += a x
__global__ void __launch_bounds__(BLK) k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
// x *= a
__global__ void __launch_bounds__(BLK) k_scal(int64_t n, double a, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) x[i] *= a;
}
// y = a - b
__global__ void __launch_bounds__(BLK) k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) y[i] = a[i] - b[i];
}
// s = w seed - D l (l may be null: s = w seed)
__global__ void __launch_bounds__(BLK) k_adj_rhs(int64_t n, double w, const double* __restrict__ seed, const double* __restrict__ D,
                                                 const double* __restrict__ l, double* __restrict__ s) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    double v = w * seed[i];
    if (l) v -= D[i] * l[i];
    s[i] = v;
}
// x = 2 a - b (linear extrapolation of the adjoint states in time)
__global__ void __launch_bounds__(BLK) k_extrap2(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                 double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) x[i] = 2.0 * a[i] - b[i];
}
__global__ void __launch_bounds__(BLK) k_zero_n(int64_t n, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) x[i] = 0.0;
}
__global__ void __launch_bounds__(BLK) k_gather(int64_t nnz, const int32_t* __restrict__ src, const double* __restrict__ in,
                                                double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k < nnz) out[k] = in[src[k]];
}
// per-block count of entries where a != b (bitwise)
__global__ void __launch_bounds__(BLK) k_neq(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                             int* __restrict__ part) {
    __shared__ int s[BLK];
    int acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK)
        acc += __double_as_longlong(a[i]) != __double_as_longlong(b[i]) ? 1 : 0;
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = BLK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

bool is_k(int k, int a, int b) { return k == a || k == b; }

// The operator list must be a flow list whose coefficients do not depend on the unknowns.
int check_flow_ops(Ctx& c, const vfvm_op* ops, int32_t nops) {
    if (!is_flow_list(ops, nops)) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow adjoint: not a flow operator list (vfvm_adjoint_forward takes the linear scalar lists)");
    for (int32_t i = 0; i < nops; ++i) {
        const vfvm_op& op = ops[i];
        if (op.kind >= 40 && op.kind < 70) continue;
        const std::string where = "flow adjoint: operator " + std::to_string(i) + ": ";
        if (is_k(op.kind, VFVM_OP_SCL_DOM_PRODUCTION, VFVM_OP_SCL_BND_PRODUCTION)) {
            if (op.a < 0 || op.a >= (int32_t)c.scl.size() || c.scl[op.a].kind != VFVM_FIELD_CONSTANT || c.scl[op.a].cnst != 0.0)
                return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, where + "production with a non-zero coefficient reads the unknown's gradient (outside the built adjoint)");
            continue;
        }
        if (is_k(op.kind, VFVM_OP_SCL_BND_DIRICHLET, VFVM_OP_SCL_BND_NEUMANN)) continue;
        if (op.a >= 0 && op.a < (int32_t)c.scl.size() && c.scl[op.a].kind == VFVM_FIELD_UNKNOWN)
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, where + "a coefficient that is an unknown makes the operator nonlinear (outside the built adjoint)");
        if (is_k(op.kind, VFVM_OP_SCL_DOM_ADVECTION, VFVM_OP_SCL_BND_ADVECTION) && op.b >= 0 && op.b < (int32_t)c.scl.size() &&
            c.scl[op.b].kind == VFVM_FIELD_UNKNOWN)
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, where + "a diffusion coefficient that is an unknown (outside the built adjoint)");
    }
    return VFVM_OK;
}

// the flow's density and viscosity must not be unknowns (flow_compile checks the rest)
int check_flow_fields(Ctx& c) {
    const FlowSys& fs = c.fs;
    if (c.scl[fs.den].kind == VFVM_FIELD_UNKNOWN || c.scl[fs.visc].kind == VFVM_FIELD_UNKNOWN)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow adjoint: density and viscosity must not be unknowns");
    for (const Ctx::VecAffine& a : c.vaff)
        if (a.vec == fs.fce && (a.scl == fs.pres))
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow adjoint: a body force that depends on the pressure is outside the built adjoint");
    return VFVM_OK;
}

// transpose of the pressure-gradient coefficient lists (host, once per mesh)
int build_gc_transpose(Ctx& c) {
    DevGeom& g = c.g;
    if (g.gct_ptr.p) return VFVM_OK;
    const int64_t N = g.N, G = g.G;
    std::vector<int32_t> ptr(N + 1, 0), slot(std::max<int64_t>(G, 1)), own(std::max<int64_t>(G, 1));
    for (int64_t q = 0; q < G; ++q) ptr[g.h_gc_vid[q] + 1]++;
    for (int64_t u = 0; u < N; ++u) ptr[u + 1] += ptr[u];
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t v = 0; v < N; ++v)
        for (int32_t q = g.h_gc_ptr[v]; q < g.h_gc_ptr[v + 1]; ++q) {
            const int32_t t = fill[g.h_gc_vid[q]]++;
            slot[t] = q;
            own[t] = (int32_t)v;
        }
    VFVM_HIP_TRY(g.gct_ptr.upload(ptr.data(), ptr.size(), c.stream));
    VFVM_HIP_TRY(g.gct_slot.upload(slot.data(), slot.size(), c.stream));
    VFVM_HIP_TRY(g.gct_own.upload(own.data(), own.size(), c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    return VFVM_OK;
}

// Work buffers of J^T products and of the outer solve
struct AdjWork {
    DevBuf<double> ec, pg, sgw, wF, wC, gu, gp1, ggp, gaP, gvelg, gfce, gfceg, atl;
    int alloc(Ctx& c) {
        const int64_t N = c.g.N, E = std::max<int64_t>(c.g.E, 1), n = c.n;
        VFVM_HIP_TRY(ec.grow(EC_ROWS * E)); VFVM_HIP_TRY(sgw.grow(3 * N));
        VFVM_HIP_TRY(pg.grow(3 * N)); VFVM_HIP_TRY(wF.grow(E)); VFVM_HIP_TRY(wC.grow(E));
        VFVM_HIP_TRY(gu.grow(3 * N)); VFVM_HIP_TRY(gp1.grow(N)); VFVM_HIP_TRY(ggp.grow(3 * N)); VFVM_HIP_TRY(gaP.grow(N));
        VFVM_HIP_TRY(gvelg.grow(9 * N)); VFVM_HIP_TRY(gfce.grow(3 * N)); VFVM_HIP_TRY(gfceg.grow(9 * N));
        VFVM_HIP_TRY(atl.grow(n));
        return VFVM_OK;
    }
};

// arguments of the context's flow system at its current state (after flow_assemble_dt)
int adj_args(Ctx& c, AdjWork& w, FlowAdjArgs& p) {
    std::memset(&p, 0, sizeof(p));
    const FlowAsm a = flow_make_args(c);
    p.g = a.g;
    p.f = a.f;
    p.pref_v = c.fs.pref_row >= 0 ? c.fs.pref_row - 3 * c.g.N : -1;
    p.nblk = (int32_t)c.sblk.size();
    p.S = 4 + p.nblk;
    if (p.nblk > 2) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow adjoint: at most two coupled scalars");
    for (int k = 0; k < p.nblk; ++k) {
        p.blk[k].prog = c.sblk[k].prog.p;
        p.blk[k].s = c.scl[c.sblk[k].unk].value.p;
    }
    p.body_blk = -1;
    for (const Ctx::VecAffine& af : c.vaff)
        if (af.vec == c.fs.fce)
            for (int k = 0; k < p.nblk; ++k)
                if (c.sblk[k].unk == af.scl) {
                    p.body_blk = k;
                    p.dfce = af.df.p;
                }
    p.rev = c.g.rev.p;
    p.gct_ptr = c.g.gct_ptr.p;
    p.gct_slot = c.g.gct_slot.p;
    p.gct_own = c.g.gct_own.p;
    p.ec = w.ec.p;
    p.pg = w.pg.p;
    p.sgw = w.sgw.p;
    p.wF = w.wF.p; p.wC = w.wC.p;
    p.gu = w.gu.p; p.gp1 = w.gp1.p; p.ggp = w.ggp.p; p.gaP = w.gaP.p;
    p.gvelg = w.gvelg.p; p.gfce = w.gfce.p; p.gfceg = w.gfceg.p;
    return VFVM_OK;
}

// y = J^T l at the state of the last assembly (A^T values gathered, pg computed)
int jt_apply(Ctx& c, AdjWork& w, FlowAdjArgs p, const double* lam, double* y) {
    if (int st = spmv_apply(c, transpose_view(c), lam, w.atl.p)) return st;
    p.lam = lam;
    p.base = w.atl.p;
    p.y = y;
    const int64_t N = c.g.N, E = c.g.E;
    if (E > 0) hipLaunchKernelGGL(k_adj_edge, dim3(grid_for(E)), dim3(BLK), 0, c.stream, p);
    hipLaunchKernelGGL(k_adj_vertex, dim3(grid_for(N * LPV)), dim3(BLK), 0, c.stream, p);
    hipLaunchKernelGGL(k_adj_final, dim3(grid_for(N * LPV)), dim3(BLK), 0, c.stream, p);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

// The system and coefficients at state x (previous state xp for a transient list): fields from
// the states, the affine closures, the flow assembly; then A^T's values and grad p - fce_c.
int load_state(Ctx& c, const std::vector<vfvm_op>& ops, double dt, const double* xp, const double* x, AdjWork& w, FlowAdjArgs& p) {
    if (xp) {
        if (int st = flow_unpack(c, xp)) return st;
        if (int st = update_prev_all(c)) return st;
    }
    if (int st = flow_unpack(c, x)) return st;
    if (int st = eval_affine_closures(c)) return st;
    if (int st = flow_assemble_dt(c, ops.data(), (int32_t)ops.size(), dt)) return st;
    if (c.nnz > 0) hipLaunchKernelGGL(k_gather, dim3(grid_for(c.nnz)), dim3(BLK), 0, c.stream, c.nnz, c.tsrc.p, c.vals.p, c.tvals.p);
    if (int st = adj_args(c, w, p)) return st;
    hipLaunchKernelGGL(k_adj_vcoeffs, dim3(grid_for(c.g.N)), dim3(BLK), 0, c.stream, p);
    if (c.g.E > 0) hipLaunchKernelGGL(k_adj_coeffs, dim3(grid_for(c.g.E)), dim3(BLK), 0, c.stream, p);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

int dot(Ctx& c, const double* a, const double* b, int64_t n, double* out) {
    VFVM_HIP_TRY(c.red.grow(DOT_BLOCKS));
    hipLaunchKernelGGL(k_dot_part, dim3(DOT_BLOCKS), dim3(BLK), 0, c.stream, n, a, b, c.red.p);
    double h[DOT_BLOCKS];
    VFVM_HIP_TRY(hipMemcpyAsync(h, c.red.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    double s = 0.0;
    for (int k = 0; k < DOT_BLOCKS; ++k) s += h[k];
    *out = s;
    return VFVM_OK;
}

// sum of x over the pressure rows but the reference one
int psum(Ctx& c, const double* x, double* out) {
    VFVM_HIP_TRY(c.red.grow(DOT_BLOCKS));
    const int S = (int)(c.n / c.g.N);
    const int64_t pref_v = c.fs.pref_row >= 0 ? c.fs.pref_row - 3 * c.g.N : -1;
    hipLaunchKernelGGL(k_psum_part, dim3(DOT_BLOCKS), dim3(BLK), 0, c.stream, c.g.N, S, pref_v, x, c.red.p);
    double h[DOT_BLOCKS];
    VFVM_HIP_TRY(hipMemcpyAsync(h, c.red.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    double s = 0.0;
    for (int k = 0; k < DOT_BLOCKS; ++k) s += h[k];
    *out = s;
    return VFVM_OK;
}

int count_neq(Ctx& c, const double* a, const double* b, int64_t n, int64_t* out) {
    VFVM_HIP_TRY(c.red_i.grow(DOT_BLOCKS));
    hipLaunchKernelGGL(k_neq, dim3(DOT_BLOCKS), dim3(BLK), 0, c.stream, n, a, b, c.red_i.p);
    int h[DOT_BLOCKS];
    VFVM_HIP_TRY(hipMemcpyAsync(h, c.red_i.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    int64_t s = 0;
    for (int k = 0; k < DOT_BLOCKS; ++k) s += h[k];
    *out = s;
    return VFVM_OK;
}

int copy_d2d(Ctx& c, double* dst, const double* src, int64_t n) {
    VFVM_HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return VFVM_OK;
}

struct OuterStats {
    int outer = 0, inner_steps = 0;
    double rel = 0.0, t_inner = 0.0;
};

// Coarse correction of the transposed solves for the reference-pressure mode.  The pressure
// is fixed by one identity row, so A 1_p = e_pref and the system restricted to the other
// unknowns is nearly singular: its near-null left vector is the constant pressure Y (the
// pressure rows but the reference one) and its near-null right vector W is the mass
// balance, the sum of the continuity rows but the reference one (the face fluxes of two
// neighbours cancel), corrected on the nodes' velocity rows (identity rows of A), so that
// W^T A is non-zero only next to the reference vertex.  A transposed solve A^T z = r then
// takes z = a W + z~, a = (Y . r) / (Y . A^T W), A^T z~ = r - a A^T W, whose right-hand side
// has no component along Y: the Krylov solve never has to build the global component a W.
struct Deflation {
    bool on = false;
    DevBuf<double> W, qW;
    double E = 0.0;
};

int deflation_setup(Ctx& c, Deflation& d) {
    d.on = false;
    const int64_t n = c.n, N = c.g.N, NP = c.g.NP;
    if (c.fs.pref_row < 0) return VFVM_OK;
    VFVM_HIP_TRY(d.W.grow(n));
    VFVM_HIP_TRY(d.qW.grow(n));
    const SysView tv = transpose_view(c);
    const int S = (int)(n / N);
    hipLaunchKernelGGL(k_cont_indicator, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, S, c.fs.pref_row - 3 * N, d.W.p);
    if (int st = spmv_apply(c, tv, d.W.p, d.qW.p)) return st;  // column sums of the continuity rows
    if (N > NP)
        hipLaunchKernelGGL(k_mass_correct, dim3(grid_for(3 * (N - NP))), dim3(BLK), 0, c.stream, N, NP, S, (const double*)d.qW.p, d.W.p);
    if (int st = spmv_apply(c, tv, d.W.p, d.qW.p)) return st;
    VFVM_HIP_TRY(hipGetLastError());
    if (int st = psum(c, d.qW.p, &d.E)) return st;
    d.on = d.E != 0.0 && std::isfinite(d.E);
    return VFVM_OK;
}

// GCR on J^T l = s, preconditioned by the transposed GMRES on A^T (flexible: each direction
// z = M^-1 r is its own inexact solve).  l holds the initial guess on entry.
int outer_solve(Ctx& c, AdjWork& w, const FlowAdjArgs& p, const vfvm_flow_adjoint_opts& o, const Deflation& dfl, const double* s,
                double* l, DevBuf<double>& ZQ, DevBuf<double>& r, DevBuf<double>& q, OuterStats& os) {
    bool setup_done = false;  // the transposed system's FP32 values and preconditioner of this state
    const int64_t n = c.n;
    const unsigned gn = grid_for(n);
    const int m = o.restart;
    os = OuterStats();
    double snorm2 = 0.0;
    if (int st = dot(c, s, s, n, &snorm2)) return st;
    const double snorm = std::sqrt(snorm2);
    if (snorm == 0.0) {
        hipLaunchKernelGGL(k_zero_n, dim3(gn), dim3(BLK), 0, c.stream, n, l);
        return VFVM_OK;
    }
    // r = s - J^T l
    if (int st = jt_apply(c, w, p, l, q.p)) return st;
    hipLaunchKernelGGL(k_sub, dim3(gn), dim3(BLK), 0, c.stream, n, s, (const double*)q.p, r.p);
    double* Z = ZQ.p;
    double* Q = ZQ.p + (size_t)m * n;
    int kept = 0;
    for (;;) {
        double rr = 0.0;
        if (int st = dot(c, r.p, r.p, n, &rr)) return st;
        os.rel = std::sqrt(rr) / snorm;
        if (!std::isfinite(os.rel)) return set_error(VFVM_ERR_LINEAR_NON_FINITE, "flow adjoint: non-finite outer residual");
        if (os.rel <= o.tol) return VFVM_OK;
        if (os.outer >= o.max_iter) break;
        // z = M^-1 r: A^T z = r from z = 0, with the coarse correction of the reference mode
        double* z = Z + (size_t)kept * n;
        double* qk = Q + (size_t)kept * n;
        VFVM_HIP_TRY(hipMemcpyAsync(c.tb.p, r.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        double a = 0.0;
        if (dfl.on) {
            double yr = 0.0;
            if (int st = psum(c, r.p, &yr)) return st;
            a = yr / dfl.E;
            hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(BLK), 0, c.stream, n, -a, (const double*)dfl.qW.p, c.tb.p);
        }
        hipLaunchKernelGGL(k_zero_n, dim3(gn), dim3(BLK), 0, c.stream, n, z);
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        auto t0 = clk::now();
        vfvm_solve_stats ss;
        const int gs = gmres_solve_inner(c, transpose_view(c), o.inner, z, ss, setup_done);
        setup_done = true;
        os.t_inner += secs(t0, clk::now());
        os.inner_steps += ss.iterations;
        if (gs != VFVM_OK && gs != VFVM_ERR_LINEAR_DID_NOT_CONVERGE) return gs;  // an inexact direction still serves
        if (dfl.on) hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(BLK), 0, c.stream, n, a, (const double*)dfl.W.p, z);
        // q_k = J^T z, orthogonalised against the kept directions (modified Gram-Schmidt)
        if (int st = jt_apply(c, w, p, z, qk)) return st;
        for (int j = 0; j < kept; ++j) {
            double beta = 0.0;
            if (int st = dot(c, qk, Q + (size_t)j * n, n, &beta)) return st;
            hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(BLK), 0, c.stream, n, -beta, (const double*)(Q + (size_t)j * n), qk);
            hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(BLK), 0, c.stream, n, -beta, (const double*)(Z + (size_t)j * n), z);
        }
        double qq = 0.0;
        if (int st = dot(c, qk, qk, n, &qq)) return st;
        if (!(qq > 0.0) || !std::isfinite(qq)) return set_error(VFVM_ERR_LINEAR_NON_FINITE, "flow adjoint: the outer direction vanished");
        const double inv = 1.0 / std::sqrt(qq);
        hipLaunchKernelGGL(k_scal, dim3(gn), dim3(BLK), 0, c.stream, n, inv, qk);
        hipLaunchKernelGGL(k_scal, dim3(gn), dim3(BLK), 0, c.stream, n, inv, z);
        double alpha = 0.0;
        if (int st = dot(c, r.p, qk, n, &alpha)) return st;
        hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(BLK), 0, c.stream, n, alpha, (const double*)z, l);
        hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(BLK), 0, c.stream, n, -alpha, (const double*)qk, r.p);
        VFVM_HIP_TRY(hipGetLastError());
        ++os.outer;
        if (++kept == m) {
            // restart: the true residual of the current iterate
            kept = 0;
            if (int st = jt_apply(c, w, p, l, q.p)) return st;
            hipLaunchKernelGGL(k_sub, dim3(gn), dim3(BLK), 0, c.stream, n, s, (const double*)q.p, r.p);
        }
    }
    char buf[160];
    snprintf(buf, sizeof(buf), "flow adjoint: outer GCR did not converge: relative residual %.3e after %d iterations", os.rel, os.outer);
    return set_error(VFVM_ERR_LINEAR_DID_NOT_CONVERGE, buf);
}

int check_state_args(Ctx& c, const vfvm_op* ops, int32_t nops, double dt) {
    if (!c.has_pattern || !c.flow) return set_error(VFVM_ERR_PATTERN_MISMATCH, "flow adjoint: no flow pattern built (vfvm_pattern_build with the flow list)");
    if (int st = check_flow_ops(c, ops, nops)) return st;
    if (int st = check_flow_fields(c)) return st;
    if (std::isnan(dt) || dt < 0.0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TIME_STEP, "dt must be >= 0");
    return VFVM_OK;
}

// device copies of host states of n entries
int upload_states(Ctx& c, const double* x, const double* xp, DevBuf<double>& dx, DevBuf<double>& dxp) {
    VFVM_HIP_TRY(dx.upload(x, c.n, c.stream));
    if (xp) VFVM_HIP_TRY(dxp.upload(xp, c.n, c.stream));
    return VFVM_OK;
}

}  // namespace

extern "C" int vfvm_flow_residual_eval(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, double dt, const double* x, const double* x_prev,
                                       double* r) {
    CTX_GUARD(ctx);
    if (!x || !r || (dt > 0.0 && !x_prev)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow residual: NULL argument");
    if (int st = check_state_args(c, ops, nops, dt)) return st;
    DevBuf<double> dx, dxp;
    if (int st = upload_states(c, x, dt > 0.0 ? x_prev : nullptr, dx, dxp)) return st;
    if (dt > 0.0) {
        if (int st = flow_unpack(c, dxp.p)) return st;
        if (int st = update_prev_all(c)) return st;
    }
    if (int st = flow_unpack(c, dx.p)) return st;
    if (int st = eval_affine_closures(c)) return st;
    if (int st = flow_assemble_dt(c, ops, nops, dt)) return st;
    DevBuf<double> ax;
    VFVM_HIP_TRY(ax.alloc(c.n));
    if (int st = spmv_apply(c, primal_view(c), dx.p, ax.p)) return st;
    hipLaunchKernelGGL(k_sub, dim3(grid_for(c.n)), dim3(BLK), 0, c.stream, c.n, (const double*)ax.p, (const double*)c.b.p, ax.p);
    VFVM_HIP_TRY(hipGetLastError());
    VFVM_HIP_TRY(hipMemcpyAsync(r, ax.p, c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_flow_jt_apply(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, double dt, const double* x, const double* x_prev,
                                  const double* lam, double* y) {
    CTX_GUARD(ctx);
    if (!x || !lam || !y || (dt > 0.0 && !x_prev)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow J^T apply: NULL argument");
    if (int st = check_state_args(c, ops, nops, dt)) return st;
    if (int st = build_transpose(c, true)) return st;
    if (int st = build_rev(c)) return st;
    if (int st = build_gc_transpose(c)) return st;
    DevBuf<double> dx, dxp, dl, dy, tmp;
    if (int st = upload_states(c, x, dt > 0.0 ? x_prev : nullptr, dx, dxp)) return st;
    const int S = (int)(c.n / c.g.N);
    VFVM_HIP_TRY(tmp.upload(lam, c.n, c.stream));
    VFVM_HIP_TRY(dl.alloc(c.n));
    VFVM_HIP_TRY(dy.alloc(c.n));
    hipLaunchKernelGGL(k_to_il, dim3(grid_for(c.n)), dim3(BLK), 0, c.stream, c.g.N, S, (const double*)tmp.p, dl.p);
    AdjWork w;
    if (int st = w.alloc(c)) return st;
    FlowAdjArgs p;
    const std::vector<vfvm_op> ol(ops, ops + nops);
    if (int st = load_state(c, ol, dt, dt > 0.0 ? dxp.p : nullptr, dx.p, w, p)) return st;
    if (int st = jt_apply(c, w, p, dl.p, dy.p)) return st;
    hipLaunchKernelGGL(k_from_il, dim3(grid_for(c.n)), dim3(BLK), 0, c.stream, c.g.N, S, (const double*)dy.p, tmp.p);
    VFVM_HIP_TRY(hipGetLastError());
    VFVM_HIP_TRY(hipMemcpyAsync(y, tmp.p, c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_flow_adjoint_forward(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts, double dt,
                                         int32_t num_ts, int32_t max_iter, double tol, double damp, int32_t checkpoint_every,
                                         const double* seed, const double* weights, vfvm_flow_adjoint_stats* stats) {
    CTX_GUARD(ctx);
    if (!opts) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow adjoint: opts is NULL");
    if (num_ts < 0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_NUM_STEPS, "flow adjoint: num_ts must be >= 0 (0: steady)");
    if (num_ts > 0 && (std::isnan(dt) || dt <= 0.0)) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TIME_STEP, "dt must be positive");
    if (max_iter <= 0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_MAX_ITER, "max_iter must be positive");
    if (std::isnan(tol) || tol <= 0.0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TOLERANCE, "tol must be positive");
    if (std::isnan(damp) || damp <= 0.0 || damp > 1.0) return set_error(VFVM_ERR_SOLVER_INVALID_DAMPING, "damp must be in (0, 1]");
    if (checkpoint_every < 1) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow adjoint: checkpoint_every must be >= 1");
    c.tape.valid = false;
    auto ts0 = clk::now();
    if (int st = vfvm_pattern_build(ctx, ops, nops, nullptr)) return st;
    if (int st = check_state_args(c, ops, nops, num_ts > 0 ? dt : 0.0)) return st;
    if (num_ts > 0 && !c.fs.transient) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "transient flow needs the velocity time operator (VEC_DOM_TIME) first");
    if (num_ts == 0 && c.fs.transient) return set_error(VFVM_ERR_OPERATOR_USE_APPLY_TIME, "a steady flow adjoint takes a list without time operators");
    AdjTape& tp = c.tape;
    tp.ops.assign(ops, ops + nops);
    tp.dt = num_ts > 0 ? dt : 0.0;
    tp.num_ts = num_ts;
    tp.every = checkpoint_every;
    tp.opts = *opts;
    tp.flow = true;
    tp.picard_max = max_iter;
    tp.picard_tol = tol;
    tp.picard_damp = damp;
    tp.picard_iterations = 0;
    tp.objective = 0.0;
    std::memset(&tp.stats, 0, sizeof(tp.stats));
    tp.seg_first = -1;
    tp.seg_count = 0;
    const int64_t n = c.n;
    DevBuf<double> dseed, x;
    if (seed) VFVM_HIP_TRY(dseed.upload(seed, n, c.stream));
    VFVM_HIP_TRY(x.alloc(n));
    VFVM_HIP_TRY(tp.final_state.alloc(n));
    vfvm_driver_stats ds;
    std::memset(&ds, 0, sizeof(ds));
    double J = 0.0;
    // the drivers' set-up (vfvm_steady_solve / vfvm_transient_solve)
    if (int st = call_update_hook(c, 0.0)) return st;
    if (int st = initial_dirichlet(c, ops, nops)) return st;
    if (num_ts > 0) {
        if (int st = call_update_hook(c, 0.0)) return st;
        if (int st = update_prev_all(c)) return st;
        const int32_t K = checkpoint_every;
        const int64_t nck = num_ts / K + 1;
        const int32_t base_last = ((num_ts - 1) / K) * K;
        VFVM_HIP_TRY(tp.ckpt.alloc((size_t)nck * n));
        VFVM_HIP_TRY(tp.seg.alloc((size_t)(std::min(K, num_ts) + 1) * n));
        if (int st = flow_pack(c, tp.ckpt.p)) return st;
        if (base_last == 0) if (int st = copy_d2d(c, tp.seg.p, tp.ckpt.p, n)) return st;
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        auto t0 = clk::now();
        tp.stats.setup_seconds = secs(ts0, t0);
        for (int32_t ts = 0; ts < num_ts; ++ts) {
            if (int st = flow_picard(c, ops, nops, *opts, max_iter, tol, damp, ds, dt, ts)) return st;
            if (int st = update_prev_all(c)) return st;
            if (int st = flow_pack(c, x.p)) return st;
            const int32_t step = ts + 1;
            if (seed) {
                double sx = 0.0;
                if (int st = dot(c, dseed.p, x.p, n, &sx)) return st;
                J += (weights ? weights[ts] : 1.0) * sx;
            }
            if (step % K == 0) if (int st = copy_d2d(c, tp.ckpt.p + (size_t)(step / K) * n, x.p, n)) return st;
            if (step >= base_last) if (int st = copy_d2d(c, tp.seg.p + (size_t)(step - base_last) * n, x.p, n)) return st;
        }
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        tp.stats.forward_seconds = secs(t0, clk::now());
        tp.seg_first = base_last;
        tp.seg_count = num_ts - base_last;
        tp.stats.checkpoints = (int32_t)nck;
    } else {
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        auto t0 = clk::now();
        tp.stats.setup_seconds = secs(ts0, t0);
        if (int st = flow_picard(c, ops, nops, *opts, max_iter, tol, damp, ds)) return st;
        if (int st = flow_pack(c, x.p)) return st;
        if (seed) if (int st = dot(c, dseed.p, x.p, n, &J)) return st;
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        tp.stats.forward_seconds = secs(t0, clk::now());
    }
    if (int st = copy_d2d(c, tp.final_state.p, x.p, n)) return st;
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    tp.stats.steps = num_ts;
    tp.stats.forward_iterations = ds.linear_iterations;
    tp.picard_iterations = ds.picard_iterations;
    tp.objective = J;
    tp.valid = true;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->base = tp.stats;
        stats->picard_iterations = ds.picard_iterations;
        stats->objective = J;
    }
    return clear_error();
}

extern "C" int vfvm_flow_adjoint_backward(vfvm_ctx* ctx, const vfvm_flow_adjoint_opts* opts, const double* seed, const double* weights,
                                          int32_t nwrt, const vfvm_adjoint_wrt* wrt, vfvm_flow_adjoint_stats* stats) {
    CTX_GUARD(ctx);
    AdjTape& tp = c.tape;
    if (!tp.valid || !tp.flow) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow adjoint: no flow forward pass recorded (call vfvm_flow_adjoint_forward)");
    if (!opts || !seed || nwrt < 0 || (nwrt > 0 && !wrt)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow adjoint: NULL or invalid argument");
    if (opts->restart < 1 || opts->restart > 64) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_RESTART, "flow adjoint: outer restart must be in [1, 64]");
    if (!(opts->tol > 0.0)) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_TOLERANCE, "flow adjoint: outer tolerance must be positive");
    if (opts->max_iter <= 0) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_MAX_OUTER, "flow adjoint: outer max_iter must be positive");
    const std::vector<vfvm_op> ops = tp.ops;
    const int32_t nops = (int32_t)ops.size();
    const DevGeom& g = c.g;
    const int64_t n = c.n, N = g.N, NP = g.NP;
    // the parameters: Dirichlet rows of coupled scalars (sclbnd) and of the velocity (vecbnd)
    for (int32_t i = 0; i < nwrt; ++i) {
        const vfvm_adjoint_wrt& w = wrt[i];
        if (!w.grad) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow adjoint: wrt[" + std::to_string(i) + "].grad is NULL");
        if (w.kind == VFVM_WRT_SCLBND) {
            if (w.id < 0 || w.id >= (int32_t)c.sbnd.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow adjoint: sclbnd id out of range");
            for (const vfvm_op& op : ops)
                if (op.kind == VFVM_OP_SCL_BND_NEUMANN && op.a == w.id)
                    return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow adjoint: gradients with respect to Neumann values are outside the built path");
        } else if (w.kind == VFVM_WRT_VECBND) {
            if (w.id < 0 || w.id >= (int32_t)c.vbndv.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow adjoint: vecbnd id out of range");
        } else {
            return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow adjoint: wrt kind must be VFVM_WRT_SCLBND or VFVM_WRT_VECBND");
        }
    }
    auto ts0 = clk::now();
    if (int st = build_transpose(c, true)) return st;
    if (int st = build_rev(c)) return st;
    if (int st = build_gc_transpose(c)) return st;
    AdjWork w;
    if (int st = w.alloc(c)) return st;
    const int m = opts->restart;
    const int S = (int)(n / N);
    DevBuf<double> dseed, lam, lamn, lam2, s, r, q, ZQ, D, lsum;
    VFVM_HIP_TRY(q.upload(seed, n, c.stream));
    VFVM_HIP_TRY(dseed.alloc(n));
    hipLaunchKernelGGL(k_to_il, dim3(grid_for(n)), dim3(BLK), 0, c.stream, N, S, (const double*)q.p, dseed.p);
    VFVM_HIP_TRY(lam.alloc(n)); VFVM_HIP_TRY(lamn.alloc(n)); VFVM_HIP_TRY(lam2.alloc(n));
    VFVM_HIP_TRY(s.alloc(n)); VFVM_HIP_TRY(r.alloc(n)); VFVM_HIP_TRY(q.alloc(n));
    VFVM_HIP_TRY(ZQ.alloc((size_t)2 * m * n));
    VFVM_HIP_TRY(D.alloc(n)); VFVM_HIP_TRY(lsum.alloc(n));
    const unsigned gn = grid_for(n);
    hipLaunchKernelGGL(k_zero_n, dim3(gn), dim3(BLK), 0, c.stream, n, lsum.p);
    hipLaunchKernelGGL(k_zero_n, dim3(gn), dim3(BLK), 0, c.stream, n, lam.p);
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    auto t0 = clk::now();
    tp.stats.setup_seconds += secs(ts0, t0);

    Deflation dfl;
    double t_rec = 0.0, t_adj = 0.0, max_rel = 0.0;
    int rec_iters = 0, rec_picard = 0, rec_steps = 0, outer_total = 0, inner_total = 0;
    int64_t mismatches = 0;
    int done = 0;
    bool have_D = false;
    // the adjoint state of the state (xp, x) of step `step` (0: steady)
    auto adjoint_state = [&](const double* xp, const double* x, int32_t step) -> int {
        FlowAdjArgs p;
        if (int st = load_state(c, ops, tp.dt, xp, x, w, p)) return st;
        if (int st = deflation_setup(c, dfl)) return st;
        if (tp.num_ts > 0 && !have_D) {
            hipLaunchKernelGGL(k_adj_time_diag, dim3(grid_for(N)), dim3(BLK), 0, c.stream, p, tp.dt, D.p);
            have_D = true;
        }
        const double wn = (tp.num_ts > 0 && weights) ? weights[step - 1] : 1.0;
        hipLaunchKernelGGL(k_adj_rhs, dim3(gn), dim3(BLK), 0, c.stream, n, wn, (const double*)dseed.p, (const double*)D.p,
                           done > 0 ? (const double*)lamn.p : nullptr, s.p);
        // initial guess: the previous adjoint state, extrapolated once two are known
        if (done >= 2) hipLaunchKernelGGL(k_extrap2, dim3(gn), dim3(BLK), 0, c.stream, n, (const double*)lamn.p, (const double*)lam2.p, lam.p);
        else if (done == 1) if (int st = copy_d2d(c, lam.p, lamn.p, n)) return st;
        VFVM_HIP_TRY(hipGetLastError());
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        auto a0 = clk::now();
        OuterStats os;
        if (int st = outer_solve(c, w, p, *opts, dfl, s.p, lam.p, ZQ, r, q, os)) return st;
        t_adj += secs(a0, clk::now());
        outer_total += os.outer;
        inner_total += os.inner_steps;
        max_rel = std::max(max_rel, os.rel);
        if (opts->on_state) opts->on_state(opts->user, step, os.outer, os.inner_steps, os.rel);
        hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(BLK), 0, c.stream, n, 1.0, (const double*)lam.p, lsum.p);
        VFVM_HIP_TRY(hipGetLastError());
        std::swap(lam2.p, lamn.p);
        std::swap(lamn.p, lam.p);
        ++done;
        return VFVM_OK;
    };

    if (tp.num_ts == 0) {
        if (int st = adjoint_state(nullptr, tp.final_state.p, 0)) return st;
    } else {
        const int32_t K = tp.every, NT = tp.num_ts;
        const int32_t s_last = (NT - 1) / K;
        const int64_t nck = NT / K + 1;
        bool verified = false;  // one replayed segment is run to its end and checked against the next checkpoint
        for (int32_t sgi = s_last; sgi >= 0; --sgi) {
            const int32_t base = sgi * K;
            const int32_t count = std::min(K, NT - base);
            if (!(tp.seg_first == base && tp.seg_count == count)) {
                // replay the segment from its checkpoint with the forward driver's steps; its last
                // state is the next checkpoint, copied instead of recomputed (except in the one
                // segment replayed to the end, whose last state is compared with the checkpoint)
                VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
                auto r0 = clk::now();
                const bool has_next = count == K && sgi + 1 < nck;
                const bool verify = has_next && !verified;
                const int32_t upto = has_next && !verify ? count - 1 : count;
                if (int st = copy_d2d(c, tp.seg.p, tp.ckpt.p + (size_t)sgi * n, n)) return st;
                if (upto < count)
                    if (int st = copy_d2d(c, tp.seg.p + (size_t)count * n, tp.ckpt.p + (size_t)(sgi + 1) * n, n)) return st;
                for (int32_t k = 1; k <= upto; ++k) {
                    if (int st = flow_unpack(c, tp.seg.p + (size_t)(k - 1) * n)) return st;
                    if (int st = update_prev_all(c)) return st;
                    vfvm_driver_stats ds;
                    std::memset(&ds, 0, sizeof(ds));
                    if (int st = flow_picard(c, ops.data(), nops, tp.opts, tp.picard_max, tp.picard_tol, tp.picard_damp, ds, tp.dt, base + k - 1)) return st;
                    if (int st = update_prev_all(c)) return st;
                    rec_iters += ds.linear_iterations;
                    rec_picard += ds.picard_iterations;
                    ++rec_steps;
                    if (int st = flow_pack(c, tp.seg.p + (size_t)k * n)) return st;
                }
                if (verify) {
                    int64_t mm = 0;
                    if (int st = count_neq(c, tp.seg.p + (size_t)count * n, tp.ckpt.p + (size_t)(sgi + 1) * n, n, &mm)) return st;
                    mismatches += mm;
                    verified = true;
                }
                VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
                t_rec += secs(r0, clk::now());
                tp.seg_first = base;
                tp.seg_count = count;
            }
            for (int32_t k = count; k >= 1; --k)
                if (int st = adjoint_state(tp.seg.p + (size_t)(k - 1) * n, tp.seg.p + (size_t)k * n, base + k)) return st;
        }
    }

    // dJ/dtheta = sum_n l^n at the rows whose b is the parameter (identity rows)
    std::vector<double> h_ls(n);
    VFVM_HIP_TRY(hipMemcpyAsync(h_ls.data(), lsum.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    for (int32_t i = 0; i < nwrt; ++i) {
        const vfvm_adjoint_wrt& wr = wrt[i];
        if (wr.kind == VFVM_WRT_SCLBND) {
            const int32_t b = c.sbnd[wr.id].bnd;
            const int64_t off = g.bnd_ptr[b], cnt = g.bnd_ptr[b + 1] - off;
            std::fill(wr.grad, wr.grad + cnt, 0.0);
            for (size_t k = 0; k < c.sblk.size(); ++k)
                for (const vfvm_op& op : c.sblk[k].ops)
                    if (op.kind == VFVM_OP_SCL_BND_DIRICHLET && op.a == wr.id)
                        for (int64_t j = 0; j < cnt; ++j) wr.grad[j] += h_ls[(int64_t)S * (NP + g.bnd_nid[off + j]) + 4 + k];
        } else {
            const int32_t b = c.vbndv[wr.id].bnd;
            const int64_t off = g.bnd_ptr[b], cnt = g.bnd_ptr[b + 1] - off;
            std::fill(wr.grad, wr.grad + 3 * cnt, 0.0);
            if (b < (int32_t)c.fs.vbnd_val.size() && c.fs.vbnd_val[b] == wr.id)
                for (int cc = 0; cc < 3; ++cc)
                    for (int64_t j = 0; j < cnt; ++j) wr.grad[cc * cnt + j] += h_ls[(int64_t)S * (NP + g.bnd_nid[off + j]) + cc];
        }
    }
    // leave the final state in the fields
    if (int st = flow_unpack(c, tp.final_state.p)) return st;
    if (tp.num_ts > 0)
        if (int st = update_prev_all(c)) return st;  // previous values as the forward pass left them
    if (int st = eval_affine_closures(c)) return st;
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    tp.stats.backward_seconds = secs(t0, clk::now());
    tp.stats.recompute_seconds = t_rec;
    tp.stats.adjoint_solve_seconds = t_adj;
    tp.stats.recompute_iterations = rec_iters;
    tp.stats.adjoint_iterations = inner_total;
    tp.stats.recomputed_steps = rec_steps;
    tp.stats.replay_mismatches = (int32_t)std::min<int64_t>(mismatches, INT32_MAX);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->base = tp.stats;
        stats->picard_iterations = tp.picard_iterations;
        stats->recompute_picard_iterations = rec_picard;
        stats->outer_iterations = outer_total;
        stats->objective = tp.objective;
        stats->max_outer_residual = max_rel;
    }
    return clear_error();
}
