// Device views of the coupled flow system shared by the flow assembly (flow.hip) and its
// adjoint (flow_adjoint.hip): the dual geometry, the flow fields with the Rhie-Chow data, the
// assembly arguments, and the face interpolations of flow_helper.rs / the flow operators.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "ctx.hpp"

namespace vfvm {

struct FlowGeom {
    int64_t N, NP, E, S, G;
    const int32_t* rp;
    const int32_t* ngh;
    const int32_t* erow;
    const double* dist;
    const double* area;
    const double* norm;  // 3E
    const double* skew;  // 3E
    const double* gw;    // 3E
    const double* vol;
    const double* cskew;  // 3N
    const int32_t* bsf_ptr;
    const int32_t* bsf_vj;
    const int32_t* bsf_vk;
    const double* bsf_area;
    const double* bsf_norm;  // 3S
    const double* bsf_wgt;   // 3S
    const int32_t* vbnd;     // vertex -> boundary (-1: point)
    const int32_t* bidx;     // index of a node in its boundary
    const int32_t* gc_ptr;
    const int32_t* gc_vid;
    const double* gc_c;  // 3G
};

struct FlowFields {
    const double* vel;    // 3N
    const double* velg;   // 9N (u_x, u_y, u_z, v_x, ...)
    const double* pres;   // N
    const double* den;    // N
    const double* deng;   // 3N
    const double* visc;   // N
    const double* viscg;  // 3N
    const double* fce;    // 3N
    const double* fceg;   // 9N
    double* rc;           // N
    double* mom;          // 3N
    double* flux;         // E
    const double* vel_prev;  // 3N (transient)
    const double* den_prev;  // N (transient)
};

// boundaries of a flow system (one domain)
constexpr int MAX_FLOW_BND = 16;

struct FlowAsm {
    FlowGeom g;
    FlowFields f;
    const int32_t* mrow;
    const int32_t* mcol;
    double* vals;
    double* b;
    const double* vbnd_val[MAX_FLOW_BND];  // per boundary: velocity values (3 x count SoA)
    int64_t vbnd_cnt[MAX_FLOW_BND];
    int64_t pref_row;  // row of the reference pressure (-1: none)
    double pref_val;
    int32_t gm, rm;  // LDS capacities of a continuity row
    int32_t transient;
    double dt;
};

// assembly arguments of the context's flow system (its pattern, fields and Rhie-Chow data)
FlowAsm flow_make_args(Ctx& c);

// skew-corrected arithmetic mean of a scalar at the face of edge (v, n)
__device__ inline double face_mean(const double* f, const double* fg, int64_t N, int64_t v, int64_t n, double sx, double sy,
                                   double sz) {
    const double m = 0.5 * (f[v] + f[n]);
    return m + 0.5 * (sx * (fg[v] + fg[n]) + sy * (fg[N + v] + fg[N + n]) + sz * (fg[2 * N + v] + fg[2 * N + n]));
}
// harmonic mean with skew correction, guarded against a non-positive result (viscosity)
__device__ inline double face_visc(const double* f, const double* fg, int64_t N, int64_t v, int64_t n, double sx, double sy,
                                   double sz) {
    const double fv = f[v], fn = f[n];
    const double m = (fv + fn != 0.0) ? 2.0 * fv * fn / (fv + fn) : 0.0;
    const double ff = m + 0.5 * (sx * (fg[v] + fg[n]) + sy * (fg[N + v] + fg[N + n]) + sz * (fg[2 * N + v] + fg[2 * N + n]));
    return ff > 0.0 ? ff : m;
}
// x^5 as powi(5) evaluates it (binary exponentiation)
__device__ inline double pow5(double x) {
    const double x2 = x * x;
    const double x4 = x2 * x2;
    return x4 * x;
}
// body force at the control-volume centroid: fce_v + grad(fce_v) . cskew_v
__device__ inline void fce_centroid(const FlowGeom& g, const FlowFields& f, int64_t v, double out[3]) {
    const int64_t N = g.N;
    const double cx = g.cskew[v], cy = g.cskew[N + v], cz = g.cskew[2 * N + v];
    const double* F = f.fceg;
    out[0] = f.fce[v] + F[v] * cx + F[N + v] * cy + F[2 * N + v] * cz;
    out[1] = f.fce[N + v] + F[3 * N + v] * cx + F[4 * N + v] * cy + F[5 * N + v] * cz;
    out[2] = f.fce[2 * N + v] + F[6 * N + v] * cx + F[7 * N + v] * cy + F[8 * N + v] * cz;
}

}  // namespace vfvm
