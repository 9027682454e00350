// Coupled velocity-pressure operators of the steady flow solver (SteadyFlow,
// src/steady/core/steady_flow.rs): the Rhie-Chow flow data (flow_helper.rs), the momentum
// rows (OpVecDom{Advection,Gradient,Diffusion,Transpose,Dilatatory,Source}), the continuity
// rows (OpSclDom{Divergence,PressureStabilizer} on the points, OpSclBnd{Divergence,
// PressureStabilizer} on the boundary nodes), the velocity Dirichlet rows (OpVecBndDirichlet)
// and the reference-pressure row (OpSclDomPressureReference), assembled on the device into
// the fixed CSR pattern of the coupled system x = [u (N), v (N), w (N), p (N)]
// (SteadyFlow::initial_start, steady_flow.rs:94-104: velocity block first, then pressure).
//
// Summation order: every matrix and b entry is the sum of its pushes in the reference's push
// order (operator order, face order, neighbour / coefficient order), with no atomics.  The
// continuity rows are wide (2-ring velocity columns, 3-ring pressure columns): one wave
// assembles one row in LDS, face by face, in phases in which each slot receives at most one
// term, or a chain of terms summed by one lane in push order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "flow.hpp"
#include "xcd.hpp"

using namespace vfvm;

namespace {

constexpr int BLK = 256;
constexpr int WAVE = 64;
constexpr int MAX_BND = MAX_FLOW_BND;
constexpr double NEG_FRAC_TWO_THREE = -2.0 / 3.0;  // opvec_dom_ddil.rs:6
inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

// flow_helper::update_flow (flow_helper.rs:3-70): rc_v = V_v / a_P with
// a_P = sum_f cond A(|Pe|) + max(F_c, 0) and the central face mass flux F_c; every vertex
__global__ void __launch_bounds__(BLK) k_flow_rc(FlowGeom g, FlowFields f) {
    const int64_t v = (int64_t)xcd_block() * BLK + threadIdx.x;
    if (v >= g.N) return;
    const int64_t N = g.N, E = g.E;
    const double uv = f.vel[v], vv = f.vel[N + v], wv = f.vel[2 * N + v];
    double coeff_p = 0.0;
    for (int32_t e = g.rp[v]; e < g.rp[v + 1]; ++e) {
        const int64_t n = g.ngh[e];
        const double face = g.area[e], dist = g.dist[e];
        const double nx = g.norm[e], ny = g.norm[E + e], nz = g.norm[2 * E + e];
        const double sx = g.skew[e], sy = g.skew[E + e], sz = g.skew[2 * E + e];
        const double den_f = face_mean(f.den, f.deng, N, v, n, sx, sy, sz);
        const double visc_f = face_visc(f.visc, f.viscg, N, v, n, sx, sy, sz);
        const double cond = visc_f * face / dist;
        const double flux = den_f * 0.5 * ((uv + f.vel[n]) * nx + (vv + f.vel[N + n]) * ny + (wv + f.vel[2 * N + n]) * nz) * face;
        const double peclet = cond != 0.0 ? flux / cond : 0.0;
        coeff_p += cond * pow5(fmax(1.0 - 0.1 * fabs(peclet), 0.0)) + fmax(flux, 0.0);
    }
    f.rc[v] = coeff_p > 0.0 ? g.vol[v] / coeff_p : 0.0;
}

// flow_helper::update_flux, first pass (flow_helper.rs:85-112): mom_v = rc_v (grad p_v - fce_c,v)
// with grad p_v = sum_m c_m p_m / V_v over grad_coeff(v)
__global__ void __launch_bounds__(BLK) k_flow_mom(FlowGeom g, FlowFields f) {
    const int64_t v = (int64_t)xcd_block() * BLK + threadIdx.x;
    if (v >= g.N) return;
    const int64_t N = g.N, G = g.G;
    const double vol = g.vol[v];
    double px = 0.0, py = 0.0, pz = 0.0;
    for (int32_t q = g.gc_ptr[v]; q < g.gc_ptr[v + 1]; ++q) {
        const double p = f.pres[g.gc_vid[q]];
        px += g.gc_c[q] * p / vol;
        py += g.gc_c[G + q] * p / vol;
        pz += g.gc_c[2 * G + q] * p / vol;
    }
    double fc[3];
    fce_centroid(g, f, v, fc);
    const double rc = f.rc[v];
    f.mom[v] = rc * (px - fc[0]);
    f.mom[N + v] = rc * (py - fc[1]);
    f.mom[2 * N + v] = rc * (pz - fc[2]);
}

// flow_helper::update_flux, second pass (flow_helper.rs:114-171): face mass flux with the
// skew-corrected face velocity and the Rhie-Chow correction; one directed edge per thread
__global__ void __launch_bounds__(BLK) k_flow_flux(FlowGeom g, FlowFields f) {
    const int64_t e = (int64_t)xcd_block() * BLK + threadIdx.x;
    if (e >= g.E) return;
    const int64_t N = g.N, E = g.E;
    const int64_t v = g.erow[e], n = g.ngh[e];
    const double face = g.area[e], dist = g.dist[e];
    const double nx = g.norm[e], ny = g.norm[E + e], nz = g.norm[2 * E + e];
    const double sx = g.skew[e], sy = g.skew[E + e], sz = g.skew[2 * E + e];
    const double den_f = face_mean(f.den, f.deng, N, v, n, sx, sy, sz);
    double uf[3];
    for (int c = 0; c < 3; ++c) {
        const double* gc = f.velg + 3 * c * N;
        uf[c] = 0.5 * (f.vel[c * N + v] + f.vel[c * N + n]) +
                0.5 * (sx * (gc[v] + gc[n]) + sy * (gc[N + v] + gc[N + n]) + sz * (gc[2 * N + v] + gc[2 * N + n]));
    }
    const double fce_e = 0.5 * (nx * (f.fce[v] + f.fce[n]) + ny * (f.fce[N + v] + f.fce[N + n]) + nz * (f.fce[2 * N + v] + f.fce[2 * N + n]));
    const double mom_e = 0.5 * (nx * (f.mom[v] + f.mom[n]) + ny * (f.mom[N + v] + f.mom[N + n]) + nz * (f.mom[2 * N + v] + f.mom[2 * N + n]));
    const double vel_rc = -0.5 * (f.rc[v] + f.rc[n]) * ((f.pres[n] - f.pres[v]) / dist - fce_e) + mom_e;
    f.flux[e] = den_f * (uf[0] * nx + uf[1] * ny + uf[2] * nz + vel_rc) * face;
}

// Momentum rows (steady_flow.rs:245-250): OpVecDomAdvection, OpVecDomGradient, OpVecDomDiffusion
// (matrix; opvec_dom_adv.rs:72-108, opvec_dom_grad.rs:56-65, opvec_dom_diff.rs:67-86) and
// OpVecDomTranspose, OpVecDomDilatatory, OpVecDomSource (b; opvec_dom_dtrn.rs:70-104,
// opvec_dom_ddil.rs:74-103, opvec_dom_src.rs:51-70) on the points; OpVecBndDirichlet rows
// (opvec_bnd_dir.rs:56-80) on the nodes.  LPR lanes per vertex, lane l holding the directed
// edges es + l and es + LPR + l; the three component rows share every coefficient.  Row
// layout: [c N + ({v} u ngh(v)) ascending | 3N + grad_coeff(v)].  The diagonal and the b sums
// walk the lanes' terms with shuffles in face order (the push order).
template <int LPR>
__global__ void __launch_bounds__(BLK) k_flow_momentum_rows(FlowAsm a) {
    const int64_t tid = (int64_t)xcd_block() * BLK + threadIdx.x;
    const int64_t v = tid / LPR;
    const int lane = (int)(tid % LPR);
    const FlowGeom& g = a.g;
    const FlowFields& f = a.f;
    if (v >= g.N) return;
    const int64_t N = g.N, E = g.E, G = g.G;
    if (v >= g.NP) {
        if (lane == 0) {
            const int32_t bnd = g.vbnd[v];
            const double* val = a.vbnd_val[bnd];
            for (int c = 0; c < 3; ++c) {
                const int64_t row = c * N + v;
                a.vals[a.mrow[row]] = 1.0;
                a.b[row] = val[c * a.vbnd_cnt[bnd] + g.bidx[v]];
            }
        }
        return;
    }
    const int32_t es = g.rp[v], ee = g.rp[v + 1], deg = ee - es;
    const int32_t gs = g.gc_ptr[v], ge = g.gc_ptr[v + 1];
    int32_t below = 0;  // neighbours below v: the diagonal's offset in the velocity part
    for (int32_t e = es; e < ee; ++e) below += g.ngh[e] < v;
    double ca[2] = {0.0, 0.0}, cd[2] = {0.0, 0.0};
    double st[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};  // transpose terms
    double sd[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};  // dilatatory terms
    const double* VG = f.velg;
    const double div_v = VG[v] + VG[4 * N + v] + VG[8 * N + v];
    for (int h = 0; h < 2; ++h) {
        const int32_t e = es + h * LPR + lane;
        if (e >= ee) continue;
        const int64_t n = g.ngh[e];
        const double face = g.area[e], dist = g.dist[e];
        const double nx = g.norm[e], ny = g.norm[E + e], nz = g.norm[2 * E + e];
        const double sx = g.skew[e], sy = g.skew[E + e], sz = g.skew[2 * E + e];
        const double visc_f = face_visc(f.visc, f.viscg, N, v, n, sx, sy, sz);
        // advection: power-law correction of the central difference plus upwinding
        const double cond = visc_f * face / dist;
        const double flux = f.flux[e];
        const double peclet = cond != 0.0 ? flux / cond : 0.0;
        const double pow_law = pow5(fmax(1.0 - 0.1 * fabs(peclet), 0.0));
        ca[h] = cond * (pow_law - 1.0) + fmax(-flux, 0.0);
        // diffusion
        cd[h] = visc_f * face / dist;
        // transpose: mean of the two vertex gradients at the face
        const double uxf = 0.5 * (VG[v] + VG[n]), uyf = 0.5 * (VG[N + v] + VG[N + n]), uzf = 0.5 * (VG[2 * N + v] + VG[2 * N + n]);
        const double vxf = 0.5 * (VG[3 * N + v] + VG[3 * N + n]), vyf = 0.5 * (VG[4 * N + v] + VG[4 * N + n]),
                     vzf = 0.5 * (VG[5 * N + v] + VG[5 * N + n]);
        const double wxf = 0.5 * (VG[6 * N + v] + VG[6 * N + n]), wyf = 0.5 * (VG[7 * N + v] + VG[7 * N + n]),
                     wzf = 0.5 * (VG[8 * N + v] + VG[8 * N + n]);
        st[h][0] = visc_f * (uxf * nx + vxf * ny + wxf * nz) * face;
        st[h][1] = visc_f * (uyf * nx + vyf * ny + wyf * nz) * face;
        st[h][2] = visc_f * (uzf * nx + vzf * ny + wzf * nz) * face;
        // dilatatory
        const double div_n = VG[n] + VG[4 * N + n] + VG[8 * N + n];
        const double lam_f = NEG_FRAC_TWO_THREE * visc_f;
        const double div_f = 0.5 * (div_v + div_n);
        const double coeff = lam_f * div_f * face;
        sd[h][0] = coeff * nx;
        sd[h][1] = coeff * ny;
        sd[h][2] = coeff * nz;
    }
    // sequential sums in face order: the diagonal (the time operator's push first in a
    // transient list, then advection pushes, then diffusion pushes) and the per-vertex sums of
    // the transpose and dilatatory operators
    double dacc = 0.0;
    if (a.transient) dacc += -f.den[v] * g.vol[v] / a.dt;  // opvec_dom_time.rs:51-56
    for (int k = 0; k < deg; ++k) {
        const double t = __shfl(k < LPR ? ca[0] : ca[1], k % LPR, LPR);
        dacc += -t;
    }
    for (int k = 0; k < deg; ++k) {
        const double t = __shfl(k < LPR ? cd[0] : cd[1], k % LPR, LPR);
        dacc += -t;
    }
    double sumt[3] = {0.0, 0.0, 0.0}, sumd[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < deg; ++k) sumt[c] += __shfl(k < LPR ? st[0][c] : st[1][c], k % LPR, LPR);
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < deg; ++k) sumd[c] += __shfl(k < LPR ? sd[0][c] : sd[1][c], k % LPR, LPR);
    for (int c = 0; c < 3; ++c) {
        const int64_t rs = a.mrow[c * N + v];
        for (int h = 0; h < 2; ++h) {
            const int32_t e = es + h * LPR + lane;
            if (e >= ee) continue;
            const int32_t off = (e - es) + (g.ngh[e] > v ? 1 : 0);
            double s = 0.0;
            s += ca[h];
            s += cd[h];
            a.vals[rs + off] = s;
        }
        if (lane == 0) a.vals[rs + below] = dacc;
        const double* cc = g.gc_c + (int64_t)c * G;
        for (int32_t q = gs + lane; q < ge; q += LPR) {
            double s = 0.0;
            s += -cc[q];
            a.vals[rs + deg + 1 + (q - gs)] = s;
        }
    }
    if (lane == 0) {
        double fc[3];
        fce_centroid(g, f, v, fc);
        const double vol = g.vol[v];
        for (int c = 0; c < 3; ++c) {
            double bb = 0.0;
            if (a.transient) bb += -f.den[v] * f.vel_prev[c * N + v] * vol / a.dt;
            bb += -sumt[c];
            bb += -sumd[c];
            bb += -fc[c] * vol;
            a.b[c * N + v] = bb;
        }
    }
}

// slot of vertex m in a sorted LDS list (binary search); the caller guarantees m is present
__device__ inline int find_slot(const int32_t* list, int len, int32_t m) {
    int lo = 0, hi = len - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (list[mid] < m) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ inline double comp3(int c, double x, double y, double z) { return c == 0 ? x : (c == 1 ? y : z); }

// Ordering point of a one-wave workgroup: a wave's LDS accesses are carried out in the order it
// issues them, so its lanes see each other's LDS writes without a barrier; this only keeps the
// compiler from moving memory accesses across the point (global loads stay in flight)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Continuity rows, one wave per vertex v: the divergence operator's velocity columns and the
// pressure stabiliser's pressure columns and b, face by face in push order, then the
// boundary sub-face terms of the divergence operator for nodes (opscl_dom_div.rs:86-150,
// opscl_dom_pstab.rs:70-148, opscl_bnd_div.rs:85-184, opscl_bnd_pstab.rs:68-151).  The two
// operators write disjoint columns, so interleaving them per face keeps every slot's order.
// The reference-pressure row is replaced by the identity (opscl_dom_pref.rs:48-63).
// LDS: us[3][gm] velocity-column sums, ps[rm] pressure-column sums, tq[3][64] chain terms,
// gcv[gm] grad_coeff(v) vertices, gsl[gm] their pressure slots, pc[rm] 3-ring vertices,
// nsl[64] velocity slots of ngh(v).
__global__ void __launch_bounds__(WAVE) k_flow_continuity_rows(FlowAsm a) {
    extern __shared__ double lds[];
    const int gm = a.gm, rm = a.rm;
    double* us = lds;
    double* ps = us + 3 * gm;
    double* tq = ps + rm;
    int32_t* gcv = reinterpret_cast<int32_t*>(tq + 3 * WAVE);
    int32_t* gsl = gcv + gm;
    int32_t* pc = gsl + gm;
    int32_t* nsl = pc + rm;
    const FlowGeom& g = a.g;
    const FlowFields& f = a.f;
    const int64_t N = g.N, E = g.E, G = g.G, S = g.S;
    const int64_t v = xcd_block();
    if (v >= N) return;
    const int lane = threadIdx.x;
    const int64_t row = 3 * N + v;
    const int64_t rs = a.mrow[row];
    if (row == a.pref_row) {
        if (lane == 0) {
            a.vals[rs] = 1.0;
            a.b[row] = a.pref_val;
        }
        return;
    }
    const int32_t gs = g.gc_ptr[v];
    const int glen = g.gc_ptr[v + 1] - gs;
    const int plen = (int)(a.mrow[row + 1] - rs) - 3 * glen;
    for (int i = lane; i < glen; i += WAVE) {
        gcv[i] = g.gc_vid[gs + i];
        us[i] = 0.0;
        us[gm + i] = 0.0;
        us[2 * gm + i] = 0.0;
    }
    for (int i = lane; i < plen; i += WAVE) {
        pc[i] = a.mcol[rs + 3 * glen + i] - (int32_t)(3 * N);
        ps[i] = 0.0;
    }
    wave_sync();
    const int32_t es = g.rp[v], ee = g.rp[v + 1];
    const int deg_v = ee - es;
    for (int i = lane; i < glen; i += WAVE) gsl[i] = find_slot(pc, plen, gcv[i]);
    if (lane < deg_v) nsl[lane] = find_slot(gcv, glen, g.ngh[es + lane]);
    const int sv = find_slot(gcv, glen, (int32_t)v);
    const int psv = find_slot(pc, plen, (int32_t)v);
    wave_sync();

    const double den_v = f.den[v];
    const double rc_v = f.rc[v];
    const double vol_v = g.vol[v];
    double fcv[3];
    fce_centroid(g, f, v, fcv);
    // Per-face data, lane q holding face q (deg_v <= 64): the face geometry, the face density,
    // the neighbour's rc / volume / list offsets, its pressure slot and the stabiliser's
    // body-force term, loaded and computed in one parallel pass and broadcast face by face
    // (every value is the expression of the sequential walk, so the sums are unchanged).
    int64_t fn = 0;
    double f_area = 0.0, f_dist = 1.0, f_nx = 0.0, f_ny = 0.0, f_nz = 0.0, f_sx = 0.0, f_sy = 0.0, f_sz = 0.0;
    double f_den = 0.0, f_rcn = 0.0, f_voln = 1.0, f_bterm = 0.0, gwx = 0.0, gwy = 0.0, gwz = 0.0;
    int32_t f_rp = 0, f_deg = 0, f_gc = 0, f_glen = 0, f_psn = 0;
    if (lane < deg_v) {
        const int64_t e = es + lane;
        fn = g.ngh[e];
        f_area = g.area[e];
        f_dist = g.dist[e];
        f_nx = g.norm[e]; f_ny = g.norm[E + e]; f_nz = g.norm[2 * E + e];
        f_sx = g.skew[e]; f_sy = g.skew[E + e]; f_sz = g.skew[2 * E + e];
        gwx = g.gw[e]; gwy = g.gw[E + e]; gwz = g.gw[2 * E + e];
        f_den = face_mean(f.den, f.deng, N, v, fn, f_sx, f_sy, f_sz);
        f_rcn = f.rc[fn];
        f_voln = g.vol[fn];
        f_rp = g.rp[fn];
        f_deg = g.rp[fn + 1] - f_rp;
        f_gc = g.gc_ptr[fn];
        f_glen = g.gc_ptr[fn + 1] - f_gc;
        f_psn = find_slot(pc, plen, (int32_t)fn);
        double fcn[3];
        fce_centroid(g, f, fn, fcn);
        double frc[3];
        for (int c = 0; c < 3; ++c)
            frc[c] = 0.5 * (rc_v + f_rcn) * 0.5 * (f.fce[c * N + v] + f.fce[c * N + fn]) - 0.5 * (rc_v * fcv[c] + f_rcn * fcn[c]);
        f_bterm = f_den * (frc[0] * f_nx + frc[1] * f_ny + frc[2] * f_nz) * f_area;
    }
    double bacc = 0.0;
    // divergence apply_time, before the stabiliser's pushes (opscl_dom_div.rs:40-66)
    if (a.transient) bacc += (den_v - f.den_prev[v]) * vol_v / a.dt;
    for (int fi = 0; fi < deg_v; ++fi) {
        const double face = __shfl(f_area, fi), dist = __shfl(f_dist, fi);
        const double nx = __shfl(f_nx, fi), ny = __shfl(f_ny, fi), nz = __shfl(f_nz, fi);
        const double sx = __shfl(f_sx, fi), sy = __shfl(f_sy, fi), sz = __shfl(f_sz, fi);
        const double den_f = __shfl(f_den, fi);
        const double rc_n = __shfl(f_rcn, fi), vol_n = __shfl(f_voln, fi);
        const int32_t ns = __shfl(f_rp, fi), deg_n = __shfl(f_deg, fi);
        const int32_t gn = __shfl(f_gc, fi), glen_n = __shfl(f_glen, fi);
        const int psn = __shfl(f_psn, fi);
        const double bterm = __shfl(f_bterm, fi);
        const int sn = nsl[fi];
        // the neighbour's lists, loaded ahead of the LDS phases: ngh(n) with its gradient
        // weights, and the first 128 entries of grad_coeff(n)
        double nw0 = 0.0, nw1 = 0.0, nw2 = 0.0;
        int32_t nnb = 0;
        if (lane < deg_n) {
            const int32_t q = ns + lane;
            nw0 = g.gw[q]; nw1 = g.gw[E + q]; nw2 = g.gw[2 * E + q];
            nnb = g.ngh[q];
        }
        int32_t gv0 = 0, gv1 = 0;
        double gx0 = 0.0, gy0 = 0.0, gz0 = 0.0, gx1 = 0.0, gy1 = 0.0, gz1 = 0.0;
        if (lane < glen_n) {
            const int64_t q = gn + lane;
            gv0 = g.gc_vid[q]; gx0 = g.gc_c[q]; gy0 = g.gc_c[G + q]; gz0 = g.gc_c[2 * G + q];
        }
        if (lane + WAVE < glen_n) {
            const int64_t q = gn + lane + WAVE;
            gv1 = g.gc_vid[q]; gx1 = g.gc_c[q]; gy1 = g.gc_c[G + q]; gz1 = g.gc_c[2 * G + q];
        }
        // ---- divergence: -rho_f vel_f . n A with the skew-corrected face velocity
        const double coeff = 0.5 * den_f * face;
        if (lane < 3) {
            const double t = -coeff * comp3(lane, nx, ny, nz);
            us[lane * gm + sn] += t;
            us[lane * gm + sv] += t;
        }
        wave_sync();
        // vertex gradient contribution: (vel_m - vel_v) terms over ngh(v)
        if (lane < deg_v) {
            const double w_skew = gwx * sx + gwy * sy + gwz * sz;
            const int sl = nsl[lane];
            for (int c = 0; c < 3; ++c) {
                const double val = -coeff * w_skew * comp3(c, nx, ny, nz);
                us[c * gm + sl] += val;
                tq[c * WAVE + lane] = val;
            }
        }
        wave_sync();
        if (lane < 3) {
            // the chain into the v slot, summed in a register in push order
            double acc = us[lane * gm + sv];
            for (int q = 0; q < deg_v; ++q) acc += -tq[lane * WAVE + q];
            us[lane * gm + sv] = acc;
        }
        wave_sync();
        // neighbour gradient contribution: (vel_m - vel_n) terms over ngh(n)
        if (lane < deg_n) {
            const double w_skew = nw0 * sx + nw1 * sy + nw2 * sz;
            const int sl = find_slot(gcv, glen, nnb);
            for (int c = 0; c < 3; ++c) {
                const double val = -coeff * w_skew * comp3(c, nx, ny, nz);
                us[c * gm + sl] += val;
                tq[c * WAVE + lane] = val;
            }
        }
        wave_sync();
        if (lane < 3) {
            double acc = us[lane * gm + sn];
            for (int q = 0; q < deg_n; ++q) acc += -tq[lane * WAVE + q];
            us[lane * gm + sn] = acc;
        }
        // ---- pressure stabiliser: Rhie-Chow pressure terms and the body-force correction
        const double cpf = den_f * 0.5 * (rc_v + rc_n) * face;
        if (lane == 0) {
            ps[psn] += cpf / dist;
            ps[psv] += -cpf / dist;
        }
        wave_sync();
        const double coeff_v = den_f * 0.5 * rc_v / vol_v * face;
        for (int i = lane; i < glen; i += WAVE) {
            const int64_t q = gs + i;
            ps[gsl[i]] += -coeff_v * (g.gc_c[q] * nx + g.gc_c[G + q] * ny + g.gc_c[2 * G + q] * nz);
        }
        wave_sync();
        const double coeff_n = den_f * 0.5 * rc_n / vol_n * face;
        if (lane < glen_n) ps[find_slot(pc, plen, gv0)] += -coeff_n * (gx0 * nx + gy0 * ny + gz0 * nz);
        if (lane + WAVE < glen_n) ps[find_slot(pc, plen, gv1)] += -coeff_n * (gx1 * nx + gy1 * ny + gz1 * nz);
        for (int i = lane + 2 * WAVE; i < glen_n; i += WAVE) {
            const int64_t q = gn + i;
            ps[find_slot(pc, plen, g.gc_vid[q])] += -coeff_n * (g.gc_c[q] * nx + g.gc_c[G + q] * ny + g.gc_c[2 * G + q] * nz);
        }
        wave_sync();
        bacc += bterm;
    }
    // boundary sub-faces of a node: -(w_m rho_m A) n_c on vertices v, j, k
    if (v >= g.NP && lane < 3) {
        for (int32_t s = g.bsf_ptr[v]; s < g.bsf_ptr[v + 1]; ++s) {
            const double area = g.bsf_area[s];
            const double n_c = comp3(lane, g.bsf_norm[s], g.bsf_norm[S + s], g.bsf_norm[2 * S + s]);
            const int32_t vj = g.bsf_vj[s], vk = g.bsf_vk[s];
            const double cv = g.bsf_wgt[s] * den_v * area;
            const double cj = g.bsf_wgt[S + s] * f.den[vj] * area;
            const double ck = g.bsf_wgt[2 * S + s] * f.den[vk] * area;
            us[lane * gm + sv] += -cv * n_c;
            us[lane * gm + find_slot(gcv, glen, vj)] += -cj * n_c;
            us[lane * gm + find_slot(gcv, glen, vk)] += -ck * n_c;
        }
    }
    wave_sync();
    for (int i = lane; i < glen; i += WAVE) {
        a.vals[rs + i] = us[i];
        a.vals[rs + glen + i] = us[gm + i];
        a.vals[rs + 2 * glen + i] = us[2 * gm + i];
    }
    for (int i = lane; i < plen; i += WAVE) a.vals[rs + 3 * glen + i] = ps[i];
    if (lane == 0) a.b[row] = bacc;
}

// OpVecBndDirichlet::apply_initial (opvec_bnd_dir.rs:34-53): node velocities from the
// boundary values
__global__ void __launch_bounds__(BLK) k_flow_dirichlet_initial(FlowAsm a, double* vel) {
    const int64_t v = a.g.NP + (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= a.g.N) return;
    const int32_t bnd = a.g.vbnd[v];
    if (bnd < 0 || bnd >= MAX_BND) return;
    const double* val = a.vbnd_val[bnd];
    if (!val) return;
    for (int c = 0; c < 3; ++c) vel[c * a.g.N + v] = val[c * a.vbnd_cnt[bnd] + a.g.bidx[v]];
}

// x = [vel | pres]
__global__ void __launch_bounds__(BLK) k_flow_pack(int64_t N, const double* __restrict__ vel, const double* __restrict__ pres,
                                                   double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < 3 * N) x[i] = vel[i];
    else if (i < 4 * N) x[i] = pres[i - 3 * N];
}
// x_damp = (1 - d) x_iter + d x_udmp (steady_base.rs:80), written to the unknowns and to xd
__global__ void __launch_bounds__(BLK) k_flow_damp(int64_t N, double d, const double* __restrict__ xi,
                                                   const double* __restrict__ xu, double* __restrict__ xd,
                                                   double* __restrict__ vel, double* __restrict__ pres) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= 4 * N) return;
    const double x = (1.0 - d) * xi[i] + d * xu[i];
    xd[i] = x;
    if (i < 3 * N) vel[i] = x;
    else pres[i - 3 * N] = x;
}

// x_damp of a coupled scalar unknown, written to its values and to xd
__global__ void __launch_bounds__(BLK) k_blk_damp(int64_t N, double d, const double* __restrict__ xi,
                                                  const double* __restrict__ xu, double* __restrict__ xd,
                                                  double* __restrict__ val) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= N) return;
    const double x = (1.0 - d) * xi[i] + d * xu[i];
    xd[i] = x;
    val[i] = x;
}

// Partials of the scaled Picard residual (scaling.rs:88-135) over vertex blocks:
// [0] sum |du|^2 over the velocity components, [1] max |vel| over both states,
// [2] max |p| over both states, [3] min p, [4] max p, [5] non-finite entries of x_new,
// [6] max |rho|
__global__ void __launch_bounds__(BLK) k_flow_stats(int64_t N, const double* __restrict__ xo, const double* __restrict__ xn,
                                                    const double* __restrict__ den, double* __restrict__ part) {
    __shared__ double s[7][BLK];
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    double r[7] = {0.0, 0.0, 0.0, INFINITY, -INFINITY, 0.0, 0.0};
    if (v < N) {
        double vo = 0.0, vn = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double a = xo[c * N + v], b = xn[c * N + v];
            const double d = fabs(b - a);
            r[0] += d * d;
            vo += a * a;
            vn += b * b;
            if (!isfinite(b)) r[5] += 1.0;
        }
        r[1] = fmax(sqrt(vo), sqrt(vn));
        const double po = xo[3 * N + v], pn = xn[3 * N + v];
        if (!isfinite(pn)) r[5] += 1.0;
        r[2] = fmax(fabs(po), fabs(pn));
        r[3] = fmin(po, pn);
        r[4] = fmax(po, pn);
        r[6] = fabs(den[v]);
    }
    for (int k = 0; k < 7; ++k) s[k][threadIdx.x] = r[k];
    __syncthreads();
    for (int off = BLK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const int t = threadIdx.x, u = t + off;
            s[0][t] += s[0][u];
            s[1][t] = fmax(s[1][t], s[1][u]);
            s[2][t] = fmax(s[2][t], s[2][u]);
            s[3][t] = fmin(s[3][t], s[3][u]);
            s[4][t] = fmax(s[4][t], s[4][u]);
            s[5][t] += s[5][u];
            s[6][t] = fmax(s[6][t], s[6][u]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) part[7 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}
// sum max(|p_new - p_old| - dead, 0)^2 (the pressure round-off dead band, scaling.rs:121-126)
__global__ void __launch_bounds__(BLK) k_flow_pres_sq(int64_t N, const double* __restrict__ po, const double* __restrict__ pn,
                                                      double dead, double* __restrict__ part) {
    __shared__ double s[BLK];
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    double r = 0.0;
    if (v < N) {
        const double d = fmax(fabs(pn[v] - po[v]) - dead, 0.0);
        r = d * d;
    }
    s[threadIdx.x] = r;
    __syncthreads();
    for (int off = BLK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// ------------------------------------------------------------------ host side

bool is_flow_kind(int k) { return k >= 40 && k < 70; }

// Validates a flow operator list: the SteadyFlow composition (steady_flow.rs:133-240) with
// one set of field ids, plus the scalar operators of coupled unknowns (heat, species:
// steady_heatflow.rs:188-330), grouped by unknown in order of first appearance.
int flow_compile(Ctx& c, const vfvm_op* ops, int32_t nops, FlowSys& fs) {
    if (!c.has_mesh) return set_error(VFVM_ERR_INVALID_ARGUMENT, "no mesh uploaded");
    if (!ops || nops <= 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "empty operator list");
    if (c.g.G == 0 || c.g.h_gc_ptr.empty())
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow operators need the mesh's pressure-gradient coefficients (vfvm_mesh_build_grad_coeff before the upload)");
    const DevGeom& g = c.g;
    const int nb = g.num_bnd;
    if (nb > MAX_BND) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "too many boundaries");
    static const int dom_order[8] = {VFVM_OP_VEC_DOM_ADVECTION, VFVM_OP_VEC_DOM_GRADIENT, VFVM_OP_VEC_DOM_DIFFUSION,
                                     VFVM_OP_VEC_DOM_TRANSPOSE, VFVM_OP_VEC_DOM_DILATATORY, VFVM_OP_VEC_DOM_SOURCE,
                                     VFVM_OP_SCL_DOM_DIVERGENCE, VFVM_OP_SCL_DOM_PRESSURE_STABILIZER};
    std::vector<vfvm_op> dom;
    std::vector<std::vector<vfvm_op>> bnd(nb);
    int npref = 0;
    vfvm_op pref{};
    fs.blk_unk.clear();
    fs.blk_ops.clear();
    for (int32_t i = 0; i < nops; ++i) {
        const vfvm_op& op = ops[i];
        if (!is_flow_kind(op.kind)) {
            // a coupled scalar unknown: its own rows after the pressure, its operators checked
            // when the pattern is built
            if (op.unk_id < 0 || op.unk_id >= (int32_t)c.scl.size() || c.scl[op.unk_id].kind != VFVM_FIELD_UNKNOWN)
                return set_error(VFVM_ERR_OPERATOR_NON_UNKNOWN, "operator " + std::to_string(i) + ": a scalar operator in a flow list must act on an unknown scalar field (OperatorNonUnknownInLinearSystem)");
            size_t k = 0;
            while (k < fs.blk_unk.size() && fs.blk_unk[k] != op.unk_id) ++k;
            if (k == fs.blk_unk.size()) {
                if (k == 2) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: at most two scalar unknowns can be coupled (six unknowns per vertex)");
                fs.blk_unk.push_back(op.unk_id);
                fs.blk_ops.emplace_back();
            }
            fs.blk_ops[k].push_back(op);
            continue;
        }
        if (npref) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "the pressure reference must come after every flow operator (it replaces its row)");
        if (op.kind < 60) {
            if (op.geom_id != 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "domain operator: only domain 0 exists");
            if (op.kind == VFVM_OP_SCL_DOM_PRESSURE_REFERENCE) { pref = op; ++npref; }
            else dom.push_back(op);
        } else {
            if (op.geom_id < 0 || op.geom_id >= nb) return set_error(VFVM_ERR_INVALID_ARGUMENT, "boundary operator: boundary id out of range");
            bnd[op.geom_id].push_back(op);
        }
    }
    fs.transient = !dom.empty() && dom[0].kind == VFVM_OP_VEC_DOM_TIME;
    vfvm_op time_op{};
    if (fs.transient) {
        time_op = dom[0];
        dom.erase(dom.begin());
    }
    if (dom.size() != 8) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: the eight domain operators of SteadyFlow are needed, once each (TransientFlow: the velocity time operator first)");
    for (int k = 0; k < 8; ++k)
        if (dom[k].kind != dom_order[k])
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: domain operators must come in SteadyFlow::assemble_matrix order (advection, gradient, diffusion, transpose, dilatatory, source, divergence, pressure stabilizer)");
    const vfvm_op& adv = dom[0];
    fs.vel = adv.unk_id;
    fs.flw = adv.c;
    if (fs.vel < 0 || fs.vel >= (int32_t)c.vec.size() || c.vec[fs.vel].kind != VFVM_FIELD_UNKNOWN)
        return set_error(VFVM_ERR_OPERATOR_NON_UNKNOWN, "flow: the velocity is not an unknown vector field (OperatorNonUnknownInLinearSystem)");
    if (fs.flw < 0 || fs.flw >= (int32_t)c.flw.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow: flow id out of range");
    const SclFlw& fl = c.flw[fs.flw];
    fs.pres = fl.pres; fs.den = fl.den; fs.visc = fl.visc; fs.fce = fl.fce;
    if (fl.vel != fs.vel) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: the operators' velocity is not the flow's velocity");
    if (c.scl[fs.pres].kind != VFVM_FIELD_UNKNOWN) return set_error(VFVM_ERR_OPERATOR_NON_UNKNOWN, "flow: the pressure is not an unknown (OperatorNonUnknownInLinearSystem)");
    for (int32_t u : fs.blk_unk)
        if (u == fs.pres) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: scalar operators cannot act on the flow's pressure");
    auto same = [&](bool ok) { return ok ? VFVM_OK : set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: every operator must name the flow's velocity, pressure, density, viscosity and body force"); };
    for (int k = 0; k < 6; ++k) if (int st = same(dom[k].unk_id == fs.vel)) return st;
    if (int st = same(adv.a == fs.den && adv.b == fs.visc)) return st;
    if (int st = same(dom[1].a == fs.pres)) return st;
    for (int k = 2; k < 5; ++k) if (int st = same(dom[k].a == fs.visc)) return st;
    if (int st = same(dom[5].a == fs.fce)) return st;
    if (int st = same(dom[6].unk_id == fs.pres && dom[6].a == fs.vel && dom[6].b == fs.den)) return st;
    if (int st = same(dom[7].unk_id == fs.pres && dom[7].a == fs.den && dom[7].b == fs.fce && dom[7].c == fs.flw)) return st;
    if (fs.transient)
        if (int st = same(time_op.unk_id == fs.vel && time_op.a == fs.den)) return st;
    fs.vbnd_val.assign(nb, -1);
    for (int b = 0; b < nb; ++b) {
        const bool has_nodes = g.bnd_ptr[b + 1] > g.bnd_ptr[b];
        if (bnd[b].empty()) {
            if (has_nodes) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: boundary " + std::to_string(b) + " has no velocity condition (pressure and traction boundaries are outside the built path)");
            continue;
        }
        if (bnd[b].size() != 3 || bnd[b][0].kind != VFVM_OP_VEC_BND_DIRICHLET || bnd[b][1].kind != VFVM_OP_SCL_BND_DIVERGENCE ||
            bnd[b][2].kind != VFVM_OP_SCL_BND_PRESSURE_STABILIZER)
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: boundary " + std::to_string(b) + " needs {VEC_BND_DIRICHLET, SCL_BND_DIVERGENCE, SCL_BND_PRESSURE_STABILIZER} in that order");
        const vfvm_op& d = bnd[b][0];
        if (int st = same(d.unk_id == fs.vel)) return st;
        if (d.a < 0 || d.a >= (int32_t)c.vbndv.size() || c.vbndv[d.a].bnd != b)
            return set_error(VFVM_ERR_INVALID_ARGUMENT, "flow: boundary velocity id out of range or on another boundary");
        fs.vbnd_val[b] = d.a;
        if (int st = same(bnd[b][1].unk_id == fs.pres && bnd[b][1].a == fs.vel && bnd[b][1].b == fs.den)) return st;
        if (int st = same(bnd[b][2].unk_id == fs.pres && bnd[b][2].a == fs.den && bnd[b][2].b == fs.fce && bnd[b][2].c == fs.flw)) return st;
    }
    fs.pref_row = -1;
    fs.pref_val = 0.0;
    if (npref) {
        if (int st = same(pref.unk_id == fs.pres)) return st;
        const int64_t nnd = g.N - g.NP;
        if (pref.a < 0 || pref.a >= nnd) return set_error(VFVM_ERR_INVALID_ARGUMENT, "pressure reference: node id out of range");
        fs.pref_row = 3 * g.N + g.NP + pref.a;
        fs.pref_val = pref.value;
    }
    return VFVM_OK;
}

size_t continuity_lds(const FlowSys& fs) {
    return sizeof(double) * (3 * (size_t)fs.gmax + fs.rmax + 3 * WAVE) + sizeof(int32_t) * (2 * (size_t)fs.gmax + fs.rmax + WAVE);
}

// gradients of the fields the flow operators read (ScalarDomain/VectorDomain::update_grad);
// a function field keeps its gradient until its values change
int flow_grads(Ctx& c, const FlowSys& fs) {
    for (int32_t id : {fs.den, fs.visc}) {
        SclDom& f = c.scl[id];
        if (f.kind == VFVM_FIELD_CONSTANT || (f.kind == VFVM_FIELD_FUNCTION && f.grad_valid)) continue;
        if (int st = launch_grad(c, f.value.p, f.grad.p)) return st;
        f.grad_valid = f.kind == VFVM_FIELD_FUNCTION;
    }
    for (int32_t id : {fs.vel, fs.fce}) {
        VecDom& f = c.vec[id];
        if (f.kind == VFVM_FIELD_CONSTANT || (f.kind == VFVM_FIELD_FUNCTION && f.grad_valid)) continue;
        if (int st = launch_grad_vec(c, f.value.p, f.grad.p)) return st;
        f.grad_valid = f.kind == VFVM_FIELD_FUNCTION;
    }
    return VFVM_OK;
}

int flow_update_impl(Ctx& c) {
    SclFlw& fl = c.flw[c.fs.flw];
    VFVM_HIP_TRY(fl.rc.alloc(c.g.N));
    VFVM_HIP_TRY(fl.mom.alloc(3 * c.g.N));
    VFVM_HIP_TRY(fl.flux.alloc(std::max<int64_t>(c.g.E, 1)));
    if (int st = flow_grads(c, c.fs)) return st;
    FlowAsm a = flow_make_args(c);
    hipLaunchKernelGGL(k_flow_rc, dim3(grid_for(c.g.N)), dim3(BLK), 0, c.stream, a.g, a.f);
    hipLaunchKernelGGL(k_flow_mom, dim3(grid_for(c.g.N)), dim3(BLK), 0, c.stream, a.g, a.f);
    if (c.g.E > 0) hipLaunchKernelGGL(k_flow_flux, dim3(grid_for(c.g.E)), dim3(BLK), 0, c.stream, a.g, a.f);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

int flow_assemble_impl(Ctx& c) {
    if (int st = flow_update_impl(c)) return st;
    FlowAsm a = flow_make_args(c);
    const int64_t N = c.g.N;
    if (c.fs.lpr == 16)
        hipLaunchKernelGGL((k_flow_momentum_rows<16>), dim3(grid_for(N * 16)), dim3(BLK), 0, c.stream, a);
    else
        hipLaunchKernelGGL((k_flow_momentum_rows<32>), dim3(grid_for(N * 32)), dim3(BLK), 0, c.stream, a);
    hipLaunchKernelGGL(k_flow_continuity_rows, dim3((unsigned)N), dim3(WAVE), continuity_lds(c.fs), c.stream, a);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

FlowAsm vfvm::flow_make_args(Ctx& c) {
    DevGeom& g = c.g;
    FlowAsm a;
    std::memset(&a, 0, sizeof(a));
    a.g = FlowGeom{g.N, g.NP, g.E, g.S, g.G, g.row_ptr.p, g.ngh.p, g.erow.p, g.dist.p, g.area.p, g.norm.p, g.skew.p, g.gw.p,
                   g.vol.p, g.cskew.p, g.bsf_ptr.p, g.bsf_vj.p, g.bsf_vk.p, g.bsf_area.p, g.bsf_norm.p, g.bsf_wgt.p,
                   g.vbnd.p, g.bidx.p, g.gc_ptr.p, g.gc_vid.p, g.gc_c.p};
    const FlowSys& fs = c.fs;
    SclFlw& fl = c.flw[fs.flw];
    a.f = FlowFields{c.vec[fs.vel].value.p, c.vec[fs.vel].grad.p, c.scl[fs.pres].value.p, c.scl[fs.den].value.p,
                     c.scl[fs.den].grad.p, c.scl[fs.visc].value.p, c.scl[fs.visc].grad.p, c.vec[fs.fce].value.p,
                     c.vec[fs.fce].grad.p, fl.rc.p, fl.mom.p, fl.flux.p, c.vec[fs.vel].value_prev.p,
                     c.scl[fs.den].value_prev.p};
    a.mrow = c.mrow.p;
    a.mcol = c.mcol.p;
    a.vals = c.vals.p;
    a.b = c.b.p;
    for (int b = 0; b < g.num_bnd && b < MAX_BND; ++b) {
        const int32_t id = b < (int)fs.vbnd_val.size() ? fs.vbnd_val[b] : -1;
        a.vbnd_val[b] = id >= 0 ? c.vbndv[id].value.p : nullptr;
        a.vbnd_cnt[b] = std::max<int64_t>(g.bnd_ptr[b + 1] - g.bnd_ptr[b], 1);
    }
    a.pref_row = fs.pref_row;
    a.pref_val = fs.pref_val;
    a.gm = fs.gmax;
    a.rm = fs.rmax;
    a.transient = fs.transient ? 1 : 0;
    a.dt = fs.dt;
    return a;
}

bool vfvm::is_flow_list(const vfvm_op* ops, int32_t nops) {
    if (!ops) return false;
    for (int32_t i = 0; i < nops; ++i)
        if (is_flow_kind(ops[i].kind)) return true;
    return false;
}

// Pattern of the coupled system (LinearBase::update_sparse of the SteadyFlow triplets):
// velocity rows of the points [c N + ({v} u ngh(v)) | 3N + grad_coeff(v)], node velocity rows
// {c N + v}, continuity rows [c N + grad_coeff(v), c = 0..2 | 3N + R3(v)] with
// R3(v) = union of grad_coeff(u) over u in {v} u ngh(v) (which holds {v} u ngh(v)), and the
// reference-pressure row {3N + v}; all ascending.
int vfvm::flow_pattern_build(Ctx& c, const vfvm_op* ops, int32_t nops) {
    if (c.g.multi)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow operators on a joined mesh (flow across interfaces) are outside the built path");
    FlowSys fs;
    if (int st = flow_compile(c, ops, nops, fs)) return st;
    const DevGeom& g = c.g;
    const int64_t N = g.N, NP = g.NP;
    const std::vector<int32_t>& rp = g.h_row_ptr;
    const std::vector<int32_t>& ng = g.h_ngh;
    const std::vector<int32_t>& gp = g.h_gc_ptr;
    const std::vector<int32_t>& gv = g.h_gc_vid;
    int32_t maxdeg = 0, maxdeg_pts = 0, gmax = 0;
    for (int64_t v = 0; v < N; ++v) {
        maxdeg = std::max(maxdeg, rp[v + 1] - rp[v]);
        if (v < NP) maxdeg_pts = std::max(maxdeg_pts, rp[v + 1] - rp[v]);
        gmax = std::max(gmax, gp[v + 1] - gp[v]);
    }
    if (maxdeg > WAVE) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: a vertex has more than 64 neighbours");
    {
        // every node needs the velocity condition of its boundary
        std::vector<uint8_t> cov(N - NP, 0);
        for (int b = 0; b < g.num_bnd; ++b)
            for (int64_t k = g.bnd_ptr[b]; k < g.bnd_ptr[b + 1]; ++k) cov[g.bnd_nid[k]] = 1;
        for (int64_t k = 0; k < N - NP; ++k)
            if (!cov[k]) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: node " + std::to_string(k) + " lies on no boundary");
    }
    fs.lpr = maxdeg_pts <= 32 ? 16 : 32;
    // 3-rings, in parallel over vertex ranges
    const unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::vector<int32_t>> r3(nth);
    std::vector<int32_t> r3len(N, 0);
    std::vector<int64_t> lo(nth + 1);
    for (unsigned t = 0; t <= nth; ++t) lo[t] = N * (int64_t)t / nth;
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; ++t)
            th.emplace_back([&, t]() {
                std::vector<int64_t> stamp(N, -1);
                std::vector<int32_t> tmp;
                for (int64_t v = lo[t]; v < lo[t + 1]; ++v) {
                    tmp.clear();
                    auto add = [&](int64_t u) {
                        for (int32_t q = gp[u]; q < gp[u + 1]; ++q) {
                            const int32_t m = gv[q];
                            if (stamp[m] != v) { stamp[m] = v; tmp.push_back(m); }
                        }
                    };
                    add(v);
                    for (int32_t e = rp[v]; e < rp[v + 1]; ++e) add(ng[e]);
                    std::sort(tmp.begin(), tmp.end());
                    r3len[v] = (int32_t)tmp.size();
                    r3[t].insert(r3[t].end(), tmp.begin(), tmp.end());
                }
            });
        for (auto& x : th) x.join();
    }
    int32_t rmax = 0;
    for (int64_t v = 0; v < N; ++v) rmax = std::max(rmax, r3len[v]);
    fs.gmax = (gmax + 1) / 2 * 2;
    fs.rmax = (rmax + 1) / 2 * 2;
    if (continuity_lds(fs) > 64 * 1024) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "flow: a continuity row does not fit in LDS");
    // row lengths
    const int64_t n = 4 * N;
    c.h_mrow.assign(n + 1, 0);
    for (int64_t v = 0; v < N; ++v) {
        const int64_t len = v < NP ? (rp[v + 1] - rp[v]) + 1 + (gp[v + 1] - gp[v]) : 1;
        for (int k = 0; k < 3; ++k) c.h_mrow[k * N + v + 1] = (int32_t)len;
    }
    for (int64_t v = 0; v < N; ++v) {
        const int64_t row = 3 * N + v;
        c.h_mrow[row + 1] = row == fs.pref_row ? 1 : 3 * (gp[v + 1] - gp[v]) + r3len[v];
    }
    fs.spmv_idx = 0;
    for (int64_t v = 0; v < N; ++v) {
        const int64_t G = gp[v + 1] - gp[v];
        const bool pref = 3 * N + v == fs.pref_row;
        fs.spmv_idx += (v < NP ? (rp[v + 1] - rp[v]) + 1 + G : 3) + (pref ? 1 : G + r3len[v]);
    }
    int64_t tot = 0;
    for (int64_t i = 0; i < n; ++i) {
        tot += c.h_mrow[i + 1];
        if (tot > INT32_MAX) return set_error(VFVM_ERR_INVALID_ARGUMENT, "pattern exceeds 32-bit indices");
        c.h_mrow[i + 1] = (int32_t)tot;
    }
    c.h_mcol.assign(tot, 0);
    std::vector<int32_t> dslot(n, -1);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; ++t)
            th.emplace_back([&, t]() {
                int64_t base = 0;  // offset of vertex v's 3-ring in r3[t]
                for (int64_t v = lo[t]; v < lo[t + 1]; ++v) {
                    const int32_t* R = r3[t].data() + base;
                    base += r3len[v];
                    const int32_t gs = gp[v], ge = gp[v + 1];
                    for (int k = 0; k < 3; ++k) {
                        const int64_t row = k * N + v;
                        int32_t w = c.h_mrow[row];
                        const int32_t off = (int32_t)(k * N);
                        if (v < NP) {
                            bool placed = false;
                            for (int32_t e = rp[v]; e < rp[v + 1]; ++e) {
                                if (!placed && ng[e] > v) { dslot[row] = w; c.h_mcol[w++] = off + (int32_t)v; placed = true; }
                                c.h_mcol[w++] = off + ng[e];
                            }
                            if (!placed) { dslot[row] = w; c.h_mcol[w++] = off + (int32_t)v; }
                            for (int32_t q = gs; q < ge; ++q) c.h_mcol[w++] = (int32_t)(3 * N) + gv[q];
                        } else {
                            dslot[row] = w;
                            c.h_mcol[w++] = off + (int32_t)v;
                        }
                    }
                    const int64_t row = 3 * N + v;
                    int32_t w = c.h_mrow[row];
                    if (row == fs.pref_row) {
                        dslot[row] = w;
                        c.h_mcol[w] = (int32_t)row;
                        continue;
                    }
                    for (int k = 0; k < 3; ++k)
                        for (int32_t q = gs; q < ge; ++q) c.h_mcol[w++] = (int32_t)(k * N) + gv[q];
                    for (int32_t i = 0; i < r3len[v]; ++i) {
                        if (R[i] == v) dslot[row] = w;
                        c.h_mcol[w++] = (int32_t)(3 * N) + R[i];
                    }
                }
            });
        for (auto& x : th) x.join();
    }
    // coupled scalar unknowns: block k holds rows (4 + k) N + v with the scalar pattern of its
    // operators ({v} u ngh(v) ascending, or {v}) shifted into its own column block
    const int nblk = (int)fs.blk_unk.size();
    std::vector<std::vector<int32_t>> b_eslot(nblk), b_dslot(nblk);
    std::vector<std::vector<uint8_t>> b_edges(nblk), b_diag(nblk);
    if (nblk > 0) {
        std::vector<int32_t> vbnd_h(N, -1);
        for (int32_t b = 0; b < g.num_bnd; ++b)
            for (int64_t k = g.bnd_ptr[b]; k < g.bnd_ptr[b + 1]; ++k) vbnd_h[NP + g.bnd_nid[k]] = b;
        c.h_mrow.resize(n + (int64_t)nblk * N + 1);
        for (int k = 0; k < nblk; ++k) {
            int32_t unk = -1;
            bool has_time = false;
            if (int st = scalar_classes(c, fs.blk_ops[k].data(), (int32_t)fs.blk_ops[k].size(), &unk, b_edges[k], b_diag[k], &has_time)) return st;
            b_eslot[k].assign(g.E, -1);
            b_dslot[k].assign(N, -1);
            const int32_t off = (int32_t)((4 + k) * N);
            for (int64_t v = 0; v < N; ++v) {
                const int cls = vbnd_h[v] < 0 ? 0 : 1 + vbnd_h[v];
                if (b_edges[k][cls]) {
                    bool placed = false;
                    for (int32_t e = rp[v]; e < rp[v + 1]; ++e) {
                        if (!placed && ng[e] > v) { b_dslot[k][v] = (int32_t)c.h_mcol.size(); c.h_mcol.push_back(off + (int32_t)v); placed = true; }
                        b_eslot[k][e] = (int32_t)c.h_mcol.size();
                        c.h_mcol.push_back(off + ng[e]);
                    }
                    if (!placed) { b_dslot[k][v] = (int32_t)c.h_mcol.size(); c.h_mcol.push_back(off + (int32_t)v); }
                } else if (b_diag[k][cls]) {
                    b_dslot[k][v] = (int32_t)c.h_mcol.size();
                    c.h_mcol.push_back(off + (int32_t)v);
                }
                if ((int64_t)c.h_mcol.size() > INT32_MAX) return set_error(VFVM_ERR_INVALID_ARGUMENT, "pattern exceeds 32-bit indices");
                const int64_t row = (4 + k) * N + v;
                c.h_mrow[row + 1] = (int32_t)c.h_mcol.size();
                fs.spmv_idx += c.h_mrow[row + 1] - c.h_mrow[row];
            }
            dslot.insert(dslot.end(), b_dslot[k].begin(), b_dslot[k].end());
        }
        tot = (int64_t)c.h_mcol.size();
    }
    const int64_t ntot = n + (int64_t)nblk * N;
    c.nnz = tot;
    c.n = ntot;
    c.unk = -1;
    hipStream_t s = c.stream;
    VFVM_HIP_TRY(c.mrow.upload(c.h_mrow.data(), c.h_mrow.size(), s));
    VFVM_HIP_TRY(c.mcol.upload(c.h_mcol.data(), c.h_mcol.size(), s));
    VFVM_HIP_TRY(c.dslot.upload(dslot.data(), dslot.size(), s));
    c.eslot.reset();
    VFVM_HIP_TRY(c.vals.alloc(std::max<int64_t>(c.nnz, 1)));
    VFVM_HIP_TRY(c.b.alloc(ntot));
    c.sblk.clear();
    c.sblk.resize(nblk);
    for (int k = 0; k < nblk; ++k) {
        ScalarBlock& blk = c.sblk[k];
        blk.unk = fs.blk_unk[k];
        blk.ops = fs.blk_ops[k];
        blk.has_edges = b_edges[k];
        blk.has_diag = b_diag[k];
        VFVM_HIP_TRY(blk.eslot.upload(b_eslot[k].data(), b_eslot[k].size(), s));
        VFVM_HIP_TRY(blk.dslot.upload(b_dslot[k].data(), b_dslot[k].size(), s));
    }
    SclFlw& fl = c.flw[fs.flw];
    VFVM_HIP_TRY(fl.rc.alloc(N));
    VFVM_HIP_TRY(fl.mom.alloc(3 * N));
    VFVM_HIP_TRY(fl.flux.alloc(std::max<int64_t>(g.E, 1)));
    VFVM_HIP_TRY(hipStreamSynchronize(s));
    c.fs = fs;
    c.flow = true;
    c.pattern_ops.assign(ops, ops + nops);
    return VFVM_OK;
}

int vfvm::flow_update(Ctx& c, int32_t flw) {
    if (flw < 0 || flw >= (int32_t)c.flw.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "update_flow: flow id out of range");
    const SclFlw& fl = c.flw[flw];
    if (c.scl[fl.pres].kind != VFVM_FIELD_UNKNOWN)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "update_flow: the flow's pressure is not an unknown (a prescribed flow's flux comes with the assembly)");
    if (c.g.G == 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "update_flow: the mesh has no pressure-gradient coefficients");
    FlowSys saved = c.fs;
    c.fs.flw = flw; c.fs.vel = fl.vel; c.fs.pres = fl.pres; c.fs.den = fl.den; c.fs.visc = fl.visc; c.fs.fce = fl.fce;
    int st = flow_update_impl(c);
    c.fs = saved;
    if (st) return st;
    return sync_err(c);
}

int vfvm::flow_assemble(Ctx& c, const vfvm_op* ops, int32_t nops) { return flow_assemble_dt(c, ops, nops, 0.0); }

int vfvm::flow_assemble_dt(Ctx& c, const vfvm_op* ops, int32_t nops, double dt) {
    if (!c.has_pattern || !c.flow) return set_error(VFVM_ERR_PATTERN_MISMATCH, "assemble: the built pattern is not a flow pattern");
    FlowSys fs;
    if (int st = flow_compile(c, ops, nops, fs)) return st;
    if (fs.vel != c.fs.vel || fs.pres != c.fs.pres || fs.flw != c.fs.flw || fs.pref_row != c.fs.pref_row || fs.blk_unk != c.fs.blk_unk)
        return set_error(VFVM_ERR_PATTERN_MISMATCH, "assemble: the flow operators do not fit the built pattern");
    if (fs.transient && !(dt > 0.0)) return set_error(VFVM_ERR_OPERATOR_USE_APPLY_TIME, "the velocity time operator needs dt > 0 (OperatorUseApplyTime)");
    c.fs.vbnd_val = fs.vbnd_val;
    c.fs.pref_val = fs.pref_val;
    c.fs.transient = fs.transient;
    c.fs.dt = fs.transient ? dt : 0.0;
    if (int st = flow_assemble_impl(c)) return st;
    // coupled scalar rows, with the face mass flux just computed (steady_heatflow.rs:466-512)
    for (size_t k = 0; k < c.sblk.size(); ++k) {
        c.sblk[k].ops = fs.blk_ops[k];
        if (int st = scalar_block_assemble(c, c.sblk[k], (int64_t)(4 + k) * c.g.N, dt, c.fs.flw)) return st;
    }
    c.has_matrix = true;
    return VFVM_OK;
}

int vfvm::flow_initial_dirichlet(Ctx& c, const vfvm_op* ops, int32_t nops) {
    FlowSys fs;
    if (int st = flow_compile(c, ops, nops, fs)) return st;
    FlowSys saved = c.fs;
    c.fs.vbnd_val = fs.vbnd_val;
    c.fs.flw = fs.flw; c.fs.vel = fs.vel; c.fs.pres = fs.pres; c.fs.den = fs.den; c.fs.visc = fs.visc; c.fs.fce = fs.fce;
    SclFlw& fl = c.flw[fs.flw];
    VFVM_HIP_TRY(fl.rc.alloc(c.g.N));
    VFVM_HIP_TRY(fl.mom.alloc(3 * c.g.N));
    FlowAsm a = flow_make_args(c);
    c.fs = saved;
    const int64_t nn = c.g.N - c.g.NP;
    if (nn > 0)
        hipLaunchKernelGGL(k_flow_dirichlet_initial, dim3(grid_for(nn)), dim3(BLK), 0, c.stream, a, c.vec[fs.vel].value.p);
    VFVM_HIP_TRY(hipGetLastError());
    // OpSclBndDirichlet::apply_initial of the coupled scalar unknowns (steady_heatflow.rs:517-528)
    for (const auto& bo : fs.blk_ops)
        if (int st = scalar_initial_dirichlet(c, bo.data(), (int32_t)bo.size())) return st;
    return sync_err(c);
}

// SteadyBase::solve for the flow system (steady_base.rs:16-135): Picard iteration with
// damping; after each solve the Rhie-Chow flow data and the matrix are rebuilt at the damped
// iterate, and the scaled residual (max over the velocity and pressure fields) is tested.
int vfvm::flow_picard(Ctx& c, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts& o, int32_t max_iter, double tol,
                      double damp, vfvm_driver_stats& ds, double dt, int ts) {
    const bool transient = dt > 0.0;
    const int64_t N = c.g.N, n = c.n;
    double* vel = c.vec[c.fs.vel].value.p;
    double* pres = c.scl[c.fs.pres].value.p;
    DevBuf<double>& xi = c.pic_xi;
    DevBuf<double>& xu = c.pic_xu;
    DevBuf<double>& xd = c.pic_xd;
    VFVM_HIP_TRY(xi.grow(n));
    VFVM_HIP_TRY(xu.grow(n));
    VFVM_HIP_TRY(xd.grow(n));
    auto t0 = clk::now();
    const double t_now = transient ? dt * (double)(ts + 1) : 0.0;  // end of the step (transient_base.rs:71)
    if (int st = flow_pack(c, xi.p)) return st;
    if (int st = call_update_hook(c, t_now)) return st;
    if (int st = flow_assemble_dt(c, ops, nops, dt)) return st;
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    ds.time_assembly += secs(t0, clk::now());
    for (int32_t iter = 0; iter < max_iter; ++iter) {
        auto ti0 = clk::now();
        VFVM_HIP_TRY(hipMemcpyAsync(xu.p, xi.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        vfvm_solve_stats st;
        if (int r = gmres_solve(c, primal_view(c), o, xu.p, st)) return r;
        ds.linear_iterations += st.iterations;
        auto ti1 = clk::now();
        ds.time_solve += secs(ti0, ti1);
        hipLaunchKernelGGL(k_flow_damp, dim3(grid_for(4 * N)), dim3(BLK), 0, c.stream, N, damp, (const double*)xi.p,
                           (const double*)xu.p, xd.p, vel, pres);
        for (size_t k = 0; k < c.sblk.size(); ++k) {
            const int64_t off = (int64_t)(4 + k) * N;
            hipLaunchKernelGGL(k_blk_damp, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, damp, (const double*)xi.p + off,
                               (const double*)xu.p + off, xd.p + off, c.scl[c.sblk[k].unk].value.p);
        }
        if (int r = call_update_hook(c, t_now)) return r;
        if (int r = flow_assemble_dt(c, ops, nops, dt)) return r;
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        ds.time_assembly += secs(ti1, clk::now());
        double res = 0.0;
        if (int r = flow_residual(c, xi.p, xd.p, &res)) return r;
        ds.residual = res;
        ds.picard_iterations += 1;
        if (!std::isfinite(res))
            return set_error(transient ? VFVM_ERR_TRANSIENT_NON_FINITE_RESIDUAL : VFVM_ERR_STEADY_NON_FINITE_RESIDUAL,
                             "non-finite Picard residual at iteration " + std::to_string(iter) + (transient ? " of step " + std::to_string(ts) : ""));
        std::swap(xi, xd);
        if (res < tol) return VFVM_OK;
    }
    char buf[160];
    snprintf(buf, sizeof(buf), "Picard iteration did not converge in %d iterations (residual %.3e)", max_iter, ds.residual);
    return set_error(transient ? VFVM_ERR_TRANSIENT_DID_NOT_CONVERGE : VFVM_ERR_STEADY_DID_NOT_CONVERGE, buf);
}

// _update_prev (transient_base.rs:63, :110): every field's previous values := current values
int vfvm::update_prev_all(Ctx& c) {
    const int64_t N = c.g.N;
    for (auto& f : c.scl)
        VFVM_HIP_TRY(hipMemcpyAsync(f.value_prev.p, f.value.p, N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    for (auto& f : c.vec)
        VFVM_HIP_TRY(hipMemcpyAsync(f.value_prev.p, f.value.p, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return sync_err(c);
}

// unknowns -> solution vector [u, v, w, p, s_1, ...] (SteadyBase::_unkn_to_solu, steady_base.rs:137-175)
int vfvm::flow_pack(Ctx& c, double* x) {
    const int64_t N = c.g.N;
    hipLaunchKernelGGL(k_flow_pack, dim3(grid_for(4 * N)), dim3(BLK), 0, c.stream, N, (const double*)c.vec[c.fs.vel].value.p,
                       (const double*)c.scl[c.fs.pres].value.p, x);
    VFVM_HIP_TRY(hipGetLastError());
    for (size_t k = 0; k < c.sblk.size(); ++k)
        VFVM_HIP_TRY(hipMemcpyAsync(x + (4 + k) * N, c.scl[c.sblk[k].unk].value.p, N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return VFVM_OK;
}

int vfvm::flow_unpack(Ctx& c, const double* x) {
    const int64_t N = c.g.N;
    VFVM_HIP_TRY(hipMemcpyAsync(c.vec[c.fs.vel].value.p, x, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    VFVM_HIP_TRY(hipMemcpyAsync(c.scl[c.fs.pres].value.p, x + 3 * N, N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    for (size_t k = 0; k < c.sblk.size(); ++k)
        VFVM_HIP_TRY(hipMemcpyAsync(c.scl[c.sblk[k].unk].value.p, x + (4 + k) * N, N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return VFVM_OK;
}

int vfvm::flow_stats(Ctx& c, const double* xo, const double* xn, double out[7]) {
    const int64_t N = c.g.N;
    const unsigned nb = grid_for(N);
    VFVM_HIP_TRY(c.red.grow(7 * (size_t)nb));
    hipLaunchKernelGGL(k_flow_stats, dim3(nb), dim3(BLK), 0, c.stream, N, xo, xn, (const double*)c.scl[c.fs.den].value.p, c.red.p);
    std::vector<double> h(7 * (size_t)nb);
    VFVM_HIP_TRY(hipMemcpyAsync(h.data(), c.red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    out[0] = 0.0; out[1] = 0.0; out[2] = 0.0; out[3] = INFINITY; out[4] = -INFINITY; out[5] = 0.0; out[6] = 0.0;
    for (unsigned k = 0; k < nb; ++k) {
        const double* p = h.data() + 7 * (size_t)k;
        out[0] += p[0];
        out[1] = std::max(out[1], p[1]);
        out[2] = std::max(out[2], p[2]);
        out[3] = std::min(out[3], p[3]);
        out[4] = std::max(out[4], p[4]);
        out[5] += p[5];
        out[6] = std::max(out[6], p[6]);
    }
    return VFVM_OK;
}

// max over the velocity, pressure and coupled scalar fields of rms(x_new - x_old) / scale
// (scaling.rs:113-135)
int vfvm::flow_residual(Ctx& c, const double* xo, const double* xn, double* res) {
    const int64_t N = c.g.N;
    if (c.scale_off) return unscaled_residual(c, xo, xn, c.n, res);
    double st[7];
    if (int r = flow_stats(c, xo, xn, st)) return r;
    if (st[5] > 0.0) { *res = NAN; return VFVM_OK; }
    const double dead = 1e-11 * st[2];  // PRES_ROUNDOFF (scaling.rs:28)
    const unsigned nb = grid_for(N);
    VFVM_HIP_TRY(c.red.grow(nb));
    hipLaunchKernelGGL(k_flow_pres_sq, dim3(nb), dim3(BLK), 0, c.stream, N, xo + 3 * N, xn + 3 * N, dead, c.red.p);
    std::vector<double> h(nb);
    VFVM_HIP_TRY(hipMemcpyAsync(h.data(), c.red.p, nb * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    double sp = 0.0;
    for (unsigned k = 0; k < nb; ++k) sp += h[k];
    const double floor = 1e-30;  // SCALE_FLOOR (scaling.rs:27)
    const double nn = (double)std::max<int64_t>(N, 1);
    const double vmax = st[1];
    // user reference scales are lower bounds (scaling.rs:95-107)
    const double scale_p = std::max(std::max(std::max(st[4] - st[3], st[6] * vmax * vmax), c.user_scale_scl(c.fs.pres)), floor);
    const double scale_v = std::max(std::max(vmax, c.user_scale_vec(c.fs.vel)), floor);
    double r = std::max(std::sqrt(sp / nn) / scale_p, std::sqrt(st[0] / nn) / scale_v);
    for (size_t k = 0; k < c.sblk.size(); ++k) {
        double rs = 0.0;
        if (int e = scalar_field_residual(c, xo + (4 + k) * N, xn + (4 + k) * N, N, c.sblk[k].unk, &rs)) return e;
        if (std::isnan(rs)) { *res = NAN; return VFVM_OK; }
        r = std::max(r, rs);
    }
    *res = r;
    return VFVM_OK;
}
