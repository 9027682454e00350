// Discrete adjoint of the linear steady / implicit-Euler map.
//
// Forward map (one linear solve per step, coefficients independent of the unknown):
//     A T^n = B T^(n-1) + f,   B = diag(-w V / dt) from the TIME operators.
// Backward: A^T l^N = seed, A^T l^n = B^T l^(n+1) (initial guess extrapolated linearly
// from l^(n+1), l^(n+2)); the states T^n are recovered by
// replaying each checkpoint segment forward with the same kernels (bit-identical, since
// no kernel on the path uses atomics).  Parameter sensitivities are linear in
//     a_e  = -sum_n l^n_v (T^n_n - T^n_v)     per directed edge e = (v, n)
//     ls_v =  sum_n l^n_v                      per vertex
//     tv_v =  sum_n l^n_v (T^n_v - T^(n-1)_v)  per vertex
// which are accumulated over the steps (the coefficients do not change with n) and
// turned into field gradients once at the end by row-gather kernels over the reverse
// edge map: no scatter, no atomics, a fixed summation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace vfvm;

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

__global__ void __launch_bounds__(BLK) k_zero(int64_t n, double* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = 0.0;
}

__global__ void __launch_bounds__(BLK) k_gather_vals(int64_t nnz, const int32_t* __restrict__ src,
                                                     const double* __restrict__ in, double* __restrict__ out) {
    int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k < nnz) out[k] = in[src[k]];
}

// x = 2 a - b: linear extrapolation in time of the adjoint state (initial guess)
__global__ void __launch_bounds__(BLK) k_extrap(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                double* __restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) x[i] = 2.0 * a[i] - b[i];
}

__global__ void __launch_bounds__(BLK) k_mul(int64_t n, const double* __restrict__ a, const double* __restrict__ x,
                                             double* __restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) y[i] = a[i] * x[i];
}

// per-block count of entries where a != b (bitwise for finite values)
__global__ void __launch_bounds__(BLK) k_count_neq(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                   int* __restrict__ part) {
    __shared__ int s[BLK];
    int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    s[threadIdx.x] = (i < n && __double_as_longlong(a[i]) != __double_as_longlong(b[i])) ? 1 : 0;
    __syncthreads();
    for (int off = BLK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// bd[v] = sum over the row class's TIME operators of -w V / dt: the derivative of b_v with
// respect to the previous state (opscl_dom_time.rs:46-58, b += -w c_prev V / dt)
__global__ void __launch_bounds__(BLK) k_time_diag(int64_t N, const int32_t* __restrict__ vbnd, const DevProgram* __restrict__ prog,
                                                   const double* __restrict__ vol, double dt, double* __restrict__ bd) {
    int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= N) return;
    const int32_t vb = vbnd[v];
    const int cls = vb < 0 ? 0 : 1 + vb;
    double acc = 0.0;
    for (int o = 0; o < prog->nops[cls]; ++o) {
        const DevOp& op = prog->ops[cls][o];
        if (op.kind == VFVM_OP_SCL_DOM_TIME || op.kind == VFVM_OP_SCL_BND_TIME) acc += -op.f_val[v] * vol[v] / dt;
    }
    bd[v] = acc;
}

// a[e] += -l_v (T_n - T_v) for every directed edge e = (v, n)
__global__ void __launch_bounds__(BLK) k_acc_edges(int64_t E, const int32_t* __restrict__ erow, const int32_t* __restrict__ ngh,
                                                   const double* __restrict__ lam, const double* __restrict__ T,
                                                   double* __restrict__ a) {
    int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (e >= E) return;
    const int64_t v = erow[e], n = ngh[e];
    a[e] += -lam[v] * (T[n] - T[v]);
}

// ls[v] += l_v; tv[v] += l_v (T_v - Tp_v) (Tp null: steady)
__global__ void __launch_bounds__(BLK) k_acc_verts(int64_t N, const double* __restrict__ lam, const double* __restrict__ T,
                                                   const double* __restrict__ Tp, double* __restrict__ ls,
                                                   double* __restrict__ tv) {
    int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= N) return;
    const double l = lam[v];
    ls[v] += l;
    if (Tp) tv[v] += l * (T[v] - Tp[v]);
}

struct VjpArgs {
    int64_t N, E;
    const int32_t* rp;
    const int32_t* erow;
    const int32_t* ngh;
    const int32_t* rev;
    const int32_t* vbnd;
    const double* area;
    const double* dist;
    const double* skew;   // 3E
    const double* gw;     // 3E
    const double* vol;
    const double* cskew;  // 3N
    const DevProgram* prog;
    const double* G;      // field values
    const double* Gg;     // field gradient (3N), zero for constant fields
    int grad_chain;       // 1: the field's gradient is computed by update_grad (non-constant)
    double dt;
    const double* a;      // edge weights
    const double* ls;
    const double* tv;
    double* pm;           // E: adjoint of the face harmonic mean
    double* gs;           // E: adjoint of the skew correction (0 where the guard dropped it)
    double* pG;           // N: direct part of dJ/dG
    double* gg;           // 3N: adjoint of grad G
    double* out;          // N
};

// Per directed edge e = (v, n): the adjoint of the face coefficient G_f of every operator of
// row v that takes G as its diffusion coefficient.  Diffusion adds c = G_f A/d
// (opscl_dom_diff.rs:78-86); power-law advection adds cond (P(Pe) - 1) + max(-F, 0) with
// cond = G_f A/d, Pe = F / cond, P = max(0, 1 - 0.1|Pe|)^5 (opscl_dom_advunity.rs:70-112),
// whose derivative in cond is P - 1 + 0.5 x^4 |Pe| with x = max(0, 1 - 0.1|Pe|).
__global__ void __launch_bounds__(BLK) k_vjp_edge(VjpArgs p) {
    int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (e >= p.E) return;
    const int64_t E = p.E, N = p.N;
    const int64_t v = p.erow[e], n = p.ngh[e];
    const int32_t vb = p.vbnd[v];
    const int cls = vb < 0 ? 0 : 1 + vb;
    const double gv = p.G[v], gn = p.G[n];
    const double sx = p.skew[e], sy = p.skew[E + e], sz = p.skew[2 * E + e];
    const double diff_m = (gv + gn != 0.0) ? 2.0 * gv * gn / (gv + gn) : 0.0;
    const double diff_s = diff_m + 0.5 * (sx * (p.Gg[v] + p.Gg[n]) + sy * (p.Gg[N + v] + p.Gg[N + n]) + sz * (p.Gg[2 * N + v] + p.Gg[2 * N + n]));
    const bool skewed = diff_s > 0.0;
    const double df = skewed ? diff_s : diff_m;
    const double ad = p.area[e] / p.dist[e];
    double dsum = 0.0;
    for (int o = 0; o < p.prog->nops[cls]; ++o) {
        const DevOp& op = p.prog->ops[cls][o];
        if (op.f_val != p.G) continue;
        if (op.kind == VFVM_OP_SCL_DOM_DIFFUSION || op.kind == VFVM_OP_SCL_BND_DIFFUSION) {
            dsum += 1.0;
        } else if ((op.kind == VFVM_OP_SCL_DOM_ADVECTION_UNITY || op.kind == VFVM_OP_SCL_BND_ADVECTION_UNITY) && op.has_diff) {
            const double cond = df * ad;
            const double pe = cond != 0.0 ? op.flux[e] / cond : 0.0;
            const double x = fmax(1.0 - 0.1 * fabs(pe), 0.0);
            const double x4 = (x * x) * (x * x);
            dsum += (x4 * x - 1.0) + 0.5 * x4 * fabs(pe);
        }
    }
    const double s = p.a[e] * ad * dsum;
    p.pm[e] = s;
    p.gs[e] = skewed ? s : 0.0;
}

// Per vertex u: the direct derivative through the harmonic means of u's faces (from both
// directed edges of each face, the reverse one through rev), through SOURCE (b += -(G_u +
// grad G_u . cskew_u) V_u, opscl_dom_src.rs:50-64) and TIME (weight G_u, opscl_dom_time.rs:46-58)
// operators of u's row; and the adjoint of grad G_u.
__global__ void __launch_bounds__(BLK) k_vjp_vertex(VjpArgs p) {
    int64_t u = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (u >= p.N) return;
    const int64_t E = p.E, N = p.N;
    const double gu = p.G[u];
    double pg = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int32_t e = p.rp[u]; e < p.rp[u + 1]; ++e) {
        const int64_t n = p.ngh[e];
        const int64_t r = p.rev[e];
        const double gn = p.G[n];
        const double sum = gu + gn;
        // d/dG_u of 2 G_u G_n / (G_u + G_n) = 2 G_n^2 / (G_u + G_n)^2, the same for both directions
        if (sum != 0.0) pg += (p.pm[e] + p.pm[r]) * (2.0 * gn * gn / (sum * sum));
        gx += 0.5 * (p.gs[e] * p.skew[e] + p.gs[r] * p.skew[r]);
        gy += 0.5 * (p.gs[e] * p.skew[E + e] + p.gs[r] * p.skew[E + r]);
        gz += 0.5 * (p.gs[e] * p.skew[2 * E + e] + p.gs[r] * p.skew[2 * E + r]);
    }
    const int32_t vb = p.vbnd[u];
    const int cls = vb < 0 ? 0 : 1 + vb;
    for (int o = 0; o < p.prog->nops[cls]; ++o) {
        const DevOp& op = p.prog->ops[cls][o];
        if (op.f_val != p.G) continue;
        if (op.kind == VFVM_OP_SCL_DOM_SOURCE || op.kind == VFVM_OP_SCL_BND_SOURCE) {
            const double w = -p.ls[u] * p.vol[u];
            pg += w;
            gx += w * p.cskew[u];
            gy += w * p.cskew[N + u];
            gz += w * p.cskew[2 * N + u];
        } else if (op.kind == VFVM_OP_SCL_DOM_TIME || op.kind == VFVM_OP_SCL_BND_TIME) {
            pg += p.tv[u] * p.vol[u] / p.dt;
        }
    }
    p.pG[u] = pg;
    p.gg[u] = gx;
    p.gg[N + u] = gy;
    p.gg[2 * N + u] = gz;
}

// dJ/dG_u = direct part + the adjoint of grad G (ScalarDomain::update_grad, scl_dom.rs:146-164:
// grad G_v = sum_loc w_(v,loc) (G_n - G_v)) gathered from u's own stencil and, through rev,
// from the stencils of u's neighbours.
__global__ void __launch_bounds__(BLK) k_vjp_final(VjpArgs p) {
    int64_t u = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (u >= p.N) return;
    const int64_t E = p.E, N = p.N;
    double acc = p.pG[u];
    if (p.grad_chain) {
        const double ux = p.gg[u], uy = p.gg[N + u], uz = p.gg[2 * N + u];
        double from_nb = 0.0, own = 0.0;
        for (int32_t e = p.rp[u]; e < p.rp[u + 1]; ++e) {
            const int64_t n = p.ngh[e];
            const int64_t r = p.rev[e];
            from_nb += p.gw[r] * p.gg[n] + p.gw[E + r] * p.gg[N + n] + p.gw[2 * E + r] * p.gg[2 * N + n];
            own += p.gw[e] * ux + p.gw[E + e] * uy + p.gw[2 * E + e] * uz;
        }
        acc += from_nb - own;
    }
    p.out[u] = acc;
}

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

bool is_kind(int k, int a, int b) { return k == a || k == b; }

// Operators whose matrix and right-hand side do not depend on the unknown's values.
int adjoint_check_ops(Ctx& c, const vfvm_op* ops, int32_t nops) {
    if (!ops || nops <= 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: empty operator list");
    if (is_flow_list(ops, nops)) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "adjoint: the flow system is nonlinear (outside the linear adjoint)");
    if (c.g.multi) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "adjoint: joined meshes (interface operators) are outside the built adjoint");
    for (int32_t i = 0; i < nops; ++i) {
        const vfvm_op& op = ops[i];
        const std::string where = "adjoint: operator " + std::to_string(i) + ": ";
        if (is_kind(op.kind, VFVM_OP_SCL_DOM_ADVECTION, VFVM_OP_SCL_BND_ADVECTION) ||
            is_kind(op.kind, VFVM_OP_SCL_DOM_TIME_PREVIOUS, VFVM_OP_SCL_BND_TIME_PREVIOUS) ||
            is_kind(op.kind, VFVM_OP_SCL_DOM_TIME_UNITY, VFVM_OP_SCL_BND_TIME_UNITY))
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, where + "no adjoint (VJP) kernels for this operator kind");
        if (is_kind(op.kind, VFVM_OP_SCL_DOM_PRODUCTION, VFVM_OP_SCL_BND_PRODUCTION)) {
            // b += -P (grad T . cskew) V reads the unknown's gradient: only an identically zero
            // production (as the heat solvers install by default) keeps the map linear
            if (op.a < 0 || op.a >= (int32_t)c.scl.size() || c.scl[op.a].kind != VFVM_FIELD_CONSTANT || c.scl[op.a].cnst != 0.0)
                return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, where + "production with a non-zero coefficient makes the map nonlinear");
            continue;
        }
        if (is_kind(op.kind, VFVM_OP_SCL_BND_DIRICHLET, VFVM_OP_SCL_BND_NEUMANN)) continue;
        if (op.a >= 0 && op.a == op.unk_id)
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, where + "a coefficient equal to the unknown makes the map nonlinear");
        if (is_kind(op.kind, VFVM_OP_SCL_DOM_ADVECTION_UNITY, VFVM_OP_SCL_BND_ADVECTION_UNITY) && op.b >= 0 &&
            op.b < (int32_t)c.flw.size() && c.flw[op.b].den == op.unk_id)
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, where + "a flow density equal to the unknown makes the map nonlinear");
    }
    return VFVM_OK;
}

}  // namespace

// CSR of A^T from the pattern of A: rows of A^T ascending in the column index of A, their
// entries ascending in the row index of A; tsrc maps each entry to its slot in A's values.
int vfvm::build_transpose(Ctx& c, bool interleaved) {
    const int64_t n = c.n, nnz = c.nnz, N = c.g.N;
    const int64_t S = interleaved && N > 0 && n % N == 0 ? n / N : 1;
    if (c.has_transpose && c.t_interleaved == (S > 1)) return VFVM_OK;
    const std::vector<int32_t>& rp = c.h_mrow;
    const std::vector<int32_t>& cl = c.h_mcol;
    // the numbering of the transposed system: identity, or vertex-interleaved (entry (k, v) of
    // the comp-major order at S v + k)
    auto perm = [&](int64_t i) -> int64_t { return S == 1 ? i : (i % N) * S + i / N; };
    c.h_trow.assign(n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) c.h_trow[perm(cl[k]) + 1]++;
    for (int64_t i = 0; i < n; ++i) c.h_trow[i + 1] += c.h_trow[i];
    c.h_tcol.assign(nnz, 0);
    std::vector<int32_t> tsrc(nnz), tdslot(n, -1), fill(c.h_trow.begin(), c.h_trow.end() - 1);
    // rows of A in increasing new number, so that each row of A^T receives its columns ascending
    for (int64_t t = 0; t < n; ++t) {
        const int64_t i = S == 1 ? t : (t % S) * N + t / S;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const int64_t j = perm(cl[k]);
            const int32_t pos = fill[j]++;
            c.h_tcol[pos] = (int32_t)t;
            tsrc[pos] = k;
            if (j == t) tdslot[j] = pos;
        }
    }
    hipStream_t s = c.stream;
    VFVM_HIP_TRY(c.trow.upload(c.h_trow.data(), c.h_trow.size(), s));
    VFVM_HIP_TRY(c.tcol.upload(c.h_tcol.data(), c.h_tcol.size(), s));
    VFVM_HIP_TRY(c.tsrc.upload(tsrc.data(), tsrc.size(), s));
    VFVM_HIP_TRY(c.tdslot.upload(tdslot.data(), tdslot.size(), s));
    VFVM_HIP_TRY(c.tvals.alloc(std::max<int64_t>(nnz, 1)));
    VFVM_HIP_TRY(c.tb.alloc(n));
    VFVM_HIP_TRY(hipStreamSynchronize(s));
    c.tilu.ready = false;
    c.has_transpose = true;
    c.t_interleaved = S > 1;
    return VFVM_OK;
}

// rev[e] for e = (v, n): the slot of v in n's (ascending) neighbour list
int vfvm::build_rev(Ctx& c) {
    DevGeom& g = c.g;
    if (g.rev.p || g.E == 0) return VFVM_OK;
    std::vector<int32_t> rev(g.E);
    for (int64_t v = 0; v < g.N; ++v)
        for (int32_t e = g.h_row_ptr[v]; e < g.h_row_ptr[v + 1]; ++e) {
            const int32_t n = g.h_ngh[e];
            auto beg = g.h_ngh.begin() + g.h_row_ptr[n], end = g.h_ngh.begin() + g.h_row_ptr[n + 1];
            auto it = std::lower_bound(beg, end, (int32_t)v);
            if (it == end || *it != (int32_t)v)
                return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: the dual's neighbour lists are not symmetric");
            rev[e] = (int32_t)(it - g.h_ngh.begin());
        }
    VFVM_HIP_TRY(g.rev.upload(rev.data(), rev.size(), c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    return VFVM_OK;
}

namespace {

int copy_d2d(Ctx& c, double* dst, const double* src, int64_t n) {
    VFVM_HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return VFVM_OK;
}

int count_neq(Ctx& c, const double* a, const double* b, int64_t n, int64_t* out) {
    const unsigned nb = grid_for(n);
    VFVM_HIP_TRY(c.red_i.grow(nb));
    hipLaunchKernelGGL(k_count_neq, dim3(nb), dim3(BLK), 0, c.stream, n, a, b, c.red_i.p);
    std::vector<int> h(nb);
    VFVM_HIP_TRY(hipMemcpyAsync(h.data(), c.red_i.p, nb * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    int64_t s = 0;
    for (int x : h) s += x;
    *out = s;
    return VFVM_OK;
}

}  // namespace

extern "C" int vfvm_adjoint_forward(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts,
                                    double dt, int32_t num_ts, int32_t checkpoint_every, vfvm_adjoint_stats* stats) {
    CTX_GUARD(ctx);
    if (!opts) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: opts is NULL");
    if (num_ts < 0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_NUM_STEPS, "adjoint: num_ts must be >= 0 (0: steady)");
    if (num_ts > 0 && (std::isnan(dt) || dt <= 0.0)) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TIME_STEP, "dt must be positive");
    if (checkpoint_every < 1) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: checkpoint_every must be >= 1");
    c.tape.valid = false;
    if (int st = adjoint_check_ops(c, ops, nops)) return st;
    auto ts0 = clk::now();
    if (int st = vfvm_pattern_build(ctx, ops, nops, nullptr)) return st;
    if (int st = initial_dirichlet(c, ops, nops)) return st;
    if (int st = solver_setup(c, primal_view(c), *opts)) return st;
    AdjTape& tp = c.tape;
    tp.ops.assign(ops, ops + nops);
    tp.dt = num_ts > 0 ? dt : 0.0;
    tp.num_ts = num_ts;
    tp.every = checkpoint_every;
    tp.opts = *opts;
    std::memset(&tp.stats, 0, sizeof(tp.stats));
    tp.seg_first = -1;
    tp.seg_count = 0;
    const int64_t n = c.n;
    SclDom& u = c.scl[c.unk];
    VFVM_HIP_TRY(c.x64.alloc(n));
    VFVM_HIP_TRY(tp.final_state.alloc(n));
    int fwd_iters = 0;
    if (num_ts > 0) {
        const int32_t K = checkpoint_every;
        const int64_t nck = num_ts / K + 1;
        const int32_t base_last = ((num_ts - 1) / K) * K;
        VFVM_HIP_TRY(tp.ckpt.alloc((size_t)nck * n));
        VFVM_HIP_TRY(tp.seg.alloc((size_t)(std::min(K, num_ts) + 1) * n));
        // the initial state is the previous state of the first step (transient_base.rs:59-62)
        if (int st = copy_d2d(c, u.value_prev.p, u.value.p, n)) return st;
        if (int st = copy_d2d(c, tp.ckpt.p, u.value.p, n)) return st;
        if (base_last == 0) if (int st = copy_d2d(c, tp.seg.p, u.value.p, n)) return st;
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        auto t0 = clk::now();
        tp.stats.setup_seconds = secs(ts0, t0);
        for (int32_t ts = 1; ts <= num_ts; ++ts) {
            if (int st = assemble_impl(c, ops, nops, dt)) return st;
            if (int st = copy_d2d(c, c.x64.p, u.value.p, n)) return st;
            vfvm_solve_stats ss;
            if (int st = gmres_solve(c, primal_view(c), *opts, c.x64.p, ss)) return st;
            fwd_iters += ss.iterations;
            if (int st = copy_d2d(c, u.value.p, c.x64.p, n)) return st;
            if (int st = copy_d2d(c, u.value_prev.p, c.x64.p, n)) return st;
            if (ts % K == 0) if (int st = copy_d2d(c, tp.ckpt.p + (size_t)(ts / K) * n, c.x64.p, n)) return st;
            if (ts >= base_last) if (int st = copy_d2d(c, tp.seg.p + (size_t)(ts - base_last) * n, c.x64.p, n)) return st;
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
        if (int st = assemble_impl(c, ops, nops, 0.0)) return st;
        if (int st = copy_d2d(c, c.x64.p, u.value.p, n)) return st;
        vfvm_solve_stats ss;
        if (int st = gmres_solve(c, primal_view(c), *opts, c.x64.p, ss)) return st;
        fwd_iters += ss.iterations;
        if (int st = copy_d2d(c, u.value.p, c.x64.p, n)) return st;
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        tp.stats.forward_seconds = secs(t0, clk::now());
    }
    if (int st = copy_d2d(c, tp.final_state.p, u.value.p, n)) return st;
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    tp.stats.steps = num_ts;
    tp.stats.forward_iterations = fwd_iters;
    tp.valid = true;
    if (stats) *stats = tp.stats;
    return clear_error();
}

extern "C" int vfvm_adjoint_backward(vfvm_ctx* ctx, const vfvm_solve_opts* opts, const double* seed, int32_t nwrt,
                                     const vfvm_adjoint_wrt* wrt, vfvm_adjoint_stats* stats) {
    CTX_GUARD(ctx);
    AdjTape& tp = c.tape;
    if (!tp.valid) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: no forward pass recorded (call vfvm_adjoint_forward)");
    if (!opts || !seed || nwrt < 0 || (nwrt > 0 && !wrt)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: NULL or invalid argument");
    const std::vector<vfvm_op>& ops = tp.ops;
    const int32_t nops = (int32_t)ops.size();
    for (int32_t i = 0; i < nwrt; ++i) {
        const vfvm_adjoint_wrt& w = wrt[i];
        if (!w.grad) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: wrt[" + std::to_string(i) + "].grad is NULL");
        if (w.kind == VFVM_WRT_SCLDOM) {
            if (w.id < 0 || w.id >= (int32_t)c.scl.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: scldom id out of range");
            if (w.id == c.unk) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: the unknown's values are the initial state (VFVM_WRT_INITIAL)");
        } else if (w.kind == VFVM_WRT_SCLBND) {
            if (w.id < 0 || w.id >= (int32_t)c.sbnd.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: sclbnd id out of range");
            for (const vfvm_op& op : ops)
                if (op.kind == VFVM_OP_SCL_BND_NEUMANN && op.a == w.id)
                    return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "adjoint: gradients with respect to Neumann values are outside the built path");
        } else if (w.kind != VFVM_WRT_INITIAL) {
            return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: unknown wrt kind");
        }
    }
    auto ts0 = clk::now();
    if (int st = build_transpose(c)) return st;
    if (int st = build_rev(c)) return st;
    if (int st = solver_setup(c, transpose_view(c), *opts)) return st;
    auto t0 = clk::now();
    tp.stats.setup_seconds += secs(ts0, t0);

    const DevGeom& g = c.g;
    const int64_t n = c.n, N = g.N, E = g.E;
    SclDom& u = c.scl[c.unk];
    DevBuf<double> lam, lamn, lam2, aE, ls, tv, bd;
    VFVM_HIP_TRY(lam.alloc(n));
    VFVM_HIP_TRY(lamn.alloc(n));
    VFVM_HIP_TRY(lam2.alloc(n));
    VFVM_HIP_TRY(aE.alloc(std::max<int64_t>(E, 1)));
    VFVM_HIP_TRY(ls.alloc(N));
    VFVM_HIP_TRY(tv.alloc(N));
    VFVM_HIP_TRY(bd.alloc(N));
    if (E == 0 || N == 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adjoint: empty mesh");
    hipLaunchKernelGGL(k_zero, dim3(grid_for(E)), dim3(BLK), 0, c.stream, E, aE.p);
    hipLaunchKernelGGL(k_zero, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, ls.p);
    hipLaunchKernelGGL(k_zero, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, tv.p);
    hipLaunchKernelGGL(k_zero, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, bd.p);
    hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, lam.p);
    VFVM_HIP_TRY(hipMemcpyAsync(c.tb.p, seed, n * sizeof(double), hipMemcpyHostToDevice, c.stream));

    // A is the same at every step of a linear map: its transpose is gathered once
    if (int st = copy_d2d(c, u.value_prev.p, tp.final_state.p, n)) return st;
    if (int st = assemble_impl(c, ops.data(), nops, tp.dt)) return st;
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(c.nnz)), dim3(BLK), 0, c.stream, c.nnz, c.tsrc.p, c.vals.p, c.tvals.p);
    if (tp.num_ts > 0)
        hipLaunchKernelGGL(k_time_diag, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, g.vbnd.p, c.prog.p, g.vol.p, tp.dt, bd.p);
    VFVM_HIP_TRY(hipGetLastError());

    double t_rec = 0.0, t_adj = 0.0;
    int rec_iters = 0, adj_iters = 0, rec_steps = 0;
    int64_t mismatches = 0;
    int done = 0;  // adjoint states computed so far
    auto adjoint_step = [&](const double* Tn, const double* Tp, bool first) -> int {
        if (!first) {
            // rhs = B^T l^(n+1); initial guess 2 l^(n+1) - l^(n+2) (l^(n+1) for the second state)
            hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, bd.p, lamn.p, c.tb.p);
            if (done >= 2) hipLaunchKernelGGL(k_extrap, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, lamn.p, lam2.p, lam.p);
            else if (int st = copy_d2d(c, lam.p, lamn.p, n)) return st;
        }
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        auto a0 = clk::now();
        vfvm_solve_stats ss;
        if (int st = gmres_solve(c, transpose_view(c), *opts, lam.p, ss)) return st;
        adj_iters += ss.iterations;
        t_adj += secs(a0, clk::now());
        hipLaunchKernelGGL(k_acc_edges, dim3(grid_for(E)), dim3(BLK), 0, c.stream, E, g.erow.p, g.ngh.p, lam.p, Tn, aE.p);
        hipLaunchKernelGGL(k_acc_verts, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, lam.p, Tn, Tp, ls.p, tv.p);
        VFVM_HIP_TRY(hipGetLastError());
        // rotate: lam2 <- l^(n+1), lamn <- l^n
        std::swap(lam2.p, lamn.p);
        std::swap(lamn.p, lam.p);
        ++done;
        return VFVM_OK;
    };

    if (tp.num_ts == 0) {
        if (int st = adjoint_step(tp.final_state.p, nullptr, true)) return st;
    } else {
        const int32_t K = tp.every, NT = tp.num_ts;
        const int32_t s_last = (NT - 1) / K;
        const int64_t nck = NT / K + 1;
        bool verified = false;  // one replayed segment is run to its end and checked against the next checkpoint
        for (int32_t sgi = s_last; sgi >= 0; --sgi) {
            const int32_t base = sgi * K;
            const int32_t count = std::min(K, NT - base);
            if (!(tp.seg_first == base && tp.seg_count == count)) {
                // replay the segment from its checkpoint, exactly as the forward pass ran it; its
                // last state is the next checkpoint, copied instead of recomputed (except in the
                // one segment replayed to the end, whose last state is compared with the checkpoint)
                VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
                auto r0 = clk::now();
                const bool has_next = count == K && sgi + 1 < nck;
                const bool verify = has_next && !verified;
                const int32_t upto = has_next && !verify ? count - 1 : count;
                if (int st = copy_d2d(c, tp.seg.p, tp.ckpt.p + (size_t)sgi * n, n)) return st;
                if (upto < count)
                    if (int st = copy_d2d(c, tp.seg.p + (size_t)count * n, tp.ckpt.p + (size_t)(sgi + 1) * n, n)) return st;
                for (int32_t k = 1; k <= upto; ++k) {
                    if (int st = copy_d2d(c, u.value_prev.p, tp.seg.p + (size_t)(k - 1) * n, n)) return st;
                    if (int st = copy_d2d(c, u.value.p, tp.seg.p + (size_t)(k - 1) * n, n)) return st;
                    if (int st = assemble_impl(c, ops.data(), nops, tp.dt)) return st;
                    if (int st = copy_d2d(c, c.x64.p, u.value.p, n)) return st;
                    vfvm_solve_stats ss;
                    if (int st = gmres_solve(c, primal_view(c), tp.opts, c.x64.p, ss)) return st;
                    rec_iters += ss.iterations;
                    ++rec_steps;
                    if (int st = copy_d2d(c, tp.seg.p + (size_t)k * n, c.x64.p, n)) return st;
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
            for (int32_t k = count; k >= 1; --k) {
                const bool first = base + k == NT;
                if (int st = adjoint_step(tp.seg.p + (size_t)k * n, tp.seg.p + (size_t)(k - 1) * n, first)) return st;
            }
        }
    }

    // vector-Jacobian products
    std::vector<double> h_ls;
    for (int32_t i = 0; i < nwrt; ++i) {
        const vfvm_adjoint_wrt& w = wrt[i];
        if (w.kind == VFVM_WRT_SCLDOM) {
            SclDom& f = c.scl[w.id];
            DevBuf<double> pm, gs, pG, gg, out;
            VFVM_HIP_TRY(pm.alloc(std::max<int64_t>(E, 1)));
            VFVM_HIP_TRY(gs.alloc(std::max<int64_t>(E, 1)));
            VFVM_HIP_TRY(pG.alloc(N));
            VFVM_HIP_TRY(gg.alloc(3 * N));
            VFVM_HIP_TRY(out.alloc(N));
            VjpArgs p;
            p.N = N; p.E = E;
            p.rp = g.row_ptr.p; p.erow = g.erow.p; p.ngh = g.ngh.p; p.rev = g.rev.p; p.vbnd = g.vbnd.p;
            p.area = g.area.p; p.dist = g.dist.p; p.skew = g.skew.p; p.gw = g.gw.p; p.vol = g.vol.p; p.cskew = g.cskew.p;
            p.prog = c.prog.p;
            p.G = f.value.p; p.Gg = f.grad.p;
            p.grad_chain = f.kind != VFVM_FIELD_CONSTANT;
            p.dt = tp.dt;
            p.a = aE.p; p.ls = ls.p; p.tv = tv.p;
            p.pm = pm.p; p.gs = gs.p; p.pG = pG.p; p.gg = gg.p; p.out = out.p;
            hipLaunchKernelGGL(k_vjp_edge, dim3(grid_for(E)), dim3(BLK), 0, c.stream, p);
            hipLaunchKernelGGL(k_vjp_vertex, dim3(grid_for(N)), dim3(BLK), 0, c.stream, p);
            hipLaunchKernelGGL(k_vjp_final, dim3(grid_for(N)), dim3(BLK), 0, c.stream, p);
            VFVM_HIP_TRY(hipGetLastError());
            VFVM_HIP_TRY(hipMemcpyAsync(w.grad, out.p, N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        } else if (w.kind == VFVM_WRT_SCLBND) {
            if (h_ls.empty()) {
                h_ls.resize(N);
                VFVM_HIP_TRY(hipMemcpyAsync(h_ls.data(), ls.p, N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
                VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
            }
            const int32_t b = c.sbnd[w.id].bnd;
            const int64_t off = g.bnd_ptr[b], cnt = g.bnd_ptr[b + 1] - off;
            std::fill(w.grad, w.grad + cnt, 0.0);
            // b_v = value (opscl_bnd_dir.rs:65-75): dJ/dvalue = sum_n l^n_v, once per operator use
            for (const vfvm_op& op : ops)
                if (op.kind == VFVM_OP_SCL_BND_DIRICHLET && op.a == w.id)
                    for (int64_t k = 0; k < cnt; ++k) w.grad[k] += h_ls[g.NP + g.bnd_nid[off + k]];
        } else {  // VFVM_WRT_INITIAL
            if (tp.num_ts == 0) {
                std::fill(w.grad, w.grad + N, 0.0);
            } else {
                // dJ/dT^0 = B^T l^1; Dirichlet nodes are overwritten before the first step
                DevBuf<double> out;
                VFVM_HIP_TRY(out.alloc(N));
                hipLaunchKernelGGL(k_mul, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, bd.p, lamn.p, out.p);
                VFVM_HIP_TRY(hipMemcpyAsync(w.grad, out.p, N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
                VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
                for (const vfvm_op& op : ops)
                    if (op.kind == VFVM_OP_SCL_BND_DIRICHLET)
                        for (int64_t k = g.bnd_ptr[op.geom_id]; k < g.bnd_ptr[op.geom_id + 1]; ++k) w.grad[g.NP + g.bnd_nid[k]] = 0.0;
            }
        }
    }
    // leave the final state in the unknown
    if (int st = copy_d2d(c, u.value.p, tp.final_state.p, n)) return st;
    if (int st = copy_d2d(c, u.value_prev.p, tp.final_state.p, n)) return st;
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    tp.stats.backward_seconds = secs(t0, clk::now());
    tp.stats.recompute_seconds = t_rec;
    tp.stats.adjoint_solve_seconds = t_adj;
    tp.stats.recompute_iterations = rec_iters;
    tp.stats.adjoint_iterations = adj_iters;
    tp.stats.recomputed_steps = rec_steps;
    tp.stats.replay_mismatches = (int32_t)std::min<int64_t>(mismatches, INT32_MAX);
    if (stats) *stats = tp.stats;
    return clear_error();
}
