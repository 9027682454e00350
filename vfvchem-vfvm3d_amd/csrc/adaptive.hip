// Adaptive time stepping over the device path (TransientAdaptive::solve_adaptive_with,
// src/transient/transient_adaptive.rs:75-350): backward Euler with the same time operators,
// a local-truncation-error estimate from the linear-extrapolation predictor through the last
// two accepted states, a PI step-size controller, rejection on error or on Picard failure,
// and steps shortened to land on the output times.  The solution vectors, the predictor and
// the error norm stay on the device; the controller (a handful of scalars) runs on the host.
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

// per-block maxima of fac |x - y_p| with the predictor y_p = x_n + ratio (x_n - x_m) (or x_n
// on the first step), and of max(|x_n|, |x|)
__global__ void __launch_bounds__(BLK) k_lte(int64_t n, const double* __restrict__ xn, const double* __restrict__ xm,
                                             const double* __restrict__ x, double ratio, double fac, int has_prev,
                                             double* __restrict__ part) {
    __shared__ double s0[BLK], s1[BLK];
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    double e = 0.0, y = 0.0;
    if (i < n) {
        const double yp = has_prev ? xn[i] + ratio * (xn[i] - xm[i]) : xn[i];
        e = fac * fabs(x[i] - yp);
        y = fmax(fabs(xn[i]), fabs(x[i]));
    }
    s0[threadIdx.x] = e;
    s1[threadIdx.x] = y;
    __syncthreads();
    for (int off = BLK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + off]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s0[0]; part[2 * blockIdx.x + 1] = s1[0]; }
}

__global__ void __launch_bounds__(BLK) k_damp_to(int64_t n, double d, const double* __restrict__ xi, const double* __restrict__ xu,
                                                 double* __restrict__ xd) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) xd[i] = (1.0 - d) * xi[i] + d * xu[i];
}

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// one unknown field of the error norm: entries [start, start + len), kind 0 scalar, 1 vector
// (3 components of `len` vertices), 2 pressure; id: its scldom / vecdom id (user scale)
struct ErrField {
    int64_t start, len;
    int kind;
    int32_t id;
};

struct System {
    Ctx& c;
    const vfvm_op* ops;
    int32_t nops;
    int64_t n;
    bool flow;
    int pack(double* x) {
        if (flow) return flow_pack(c, x);
        VFVM_HIP_TRY(hipMemcpyAsync(x, c.scl[c.unk].value.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        return VFVM_OK;
    }
    int unpack(const double* x) {
        if (flow) return flow_unpack(c, x);
        VFVM_HIP_TRY(hipMemcpyAsync(c.scl[c.unk].value.p, x, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        return VFVM_OK;
    }
    int assemble(double dt) { return flow ? flow_assemble_dt(c, ops, nops, dt) : assemble_impl(c, ops, nops, dt); }
    int residual(const double* xo, const double* xn, double* res) {
        return flow ? flow_residual(c, xo, xn, res) : scalar_picard_residual(c, xo, xn, res);
    }
};

// max over fields of max_i |LTE_i| / (atol + round-off + rtol s_f)  (transient_adaptive.rs:247-262)
int error_norm(System& sy, const std::vector<ErrField>& fields, const double* xn, const double* xm, const double* x, double ratio,
               double fac, bool has_prev, double atol, double rtol, double* err) {
    Ctx& c = sy.c;
    double e = 0.0;
    double st[7] = {0, 0, 0, 0, 0, 0, 0};
    if (sy.flow)
        if (int r = flow_stats(c, xn, x, st)) return r;
    for (const ErrField& f : fields) {
        const int64_t cnt = f.kind == 1 ? 3 * f.len : f.len;
        const unsigned nb = grid_for(cnt);
        VFVM_HIP_TRY(c.red.grow(2 * (size_t)nb));
        double* part = c.red.p;
        hipLaunchKernelGGL(k_lte, dim3(nb), dim3(BLK), 0, c.stream, cnt, xn + f.start, xm ? xm + f.start : xn + f.start,
                           x + f.start, ratio, fac, has_prev ? 1 : 0, part);
        std::vector<double> h(2 * (size_t)nb);
        VFVM_HIP_TRY(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        double emax = 0.0, ymax = 0.0;
        for (unsigned k = 0; k < nb; ++k) { emax = std::max(emax, h[2 * k]); ymax = std::max(ymax, h[2 * k + 1]); }
        // field scale (scaling.rs:88-111; a user scale is a lower bound) and the pressure
        // round-off (scaling.rs:30-41); with scale_off every block is measured against its own
        // max(|y_n|, |y|) with no round-off term (transient_adaptive.rs:257-270)
        double scale = ymax, round = 0.0;
        if (!c.scale_off) {
            if (f.kind == 0) scale = std::max(ymax, c.user_scale_scl(f.id));
            if (f.kind == 1) scale = std::max(st[1], c.user_scale_vec(f.id));
            if (f.kind == 2) {
                scale = std::max(std::max(st[4] - st[3], st[6] * st[1] * st[1]), c.user_scale_scl(f.id));
                round = 1e-11 * st[2];
            }
        }
        scale = std::max(scale, 1e-30);
        e = std::max(e, emax / (atol + round + rtol * scale));
    }
    *err = e;
    return VFVM_OK;
}

}  // namespace

extern "C" void vfvm_adaptive_opts_default(double dt_init, double t_end, vfvm_adaptive_opts* o) {
    // AdaptiveOptions::new (transient_adaptive.rs:62-73)
    std::memset(o, 0, sizeof(*o));
    o->t_end = t_end;
    o->dt_init = dt_init;
    o->dt_min = 1e-12 * t_end;
    o->dt_max = t_end;
    o->rtol = 1e-3;
    o->atol = 1e-12;
    o->safety = 0.9;
    o->fac_min = 0.2;
    o->fac_max = 2.0;
    o->fac_fail = 0.25;
    o->max_reject = 30;
    o->max_iter = 30;
    o->tol = 1e-6;
    o->damp = 1.0;
    o->fixed = 0;
    o->num_out = -1;  // 11 evenly spaced output times from 0 to t_end
    o->out_times = nullptr;
    o->num_exclude = 0;
    o->scl_exclude = nullptr;
}

extern "C" int vfvm_transient_adaptive(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* lin,
                                       const vfvm_adaptive_opts* opt, vfvm_step_hook hook, void* user, double* steps_out,
                                       int32_t max_steps, vfvm_adaptive_stats* stats) {
    CTX_GUARD(ctx);
    if (!opt || !lin) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adaptive: options are NULL");
    // transient_adaptive.rs:83-89
    if (!(opt->t_end > 0.0)) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TIME_STEP, "adaptive: t_end must be positive");
    if (!(opt->dt_init > 0.0) || !(opt->dt_min > 0.0) || !(opt->dt_max >= opt->dt_min))
        return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TIME_STEP, "adaptive: dt_init, dt_min must be positive and dt_max >= dt_min");
    if (opt->max_iter <= 0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_MAX_ITER, "adaptive: max_iter must be positive");
    if (!(opt->tol > 0.0)) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TOLERANCE, "adaptive: tol must be positive");
    if (!(opt->damp > 0.0 && opt->damp <= 1.0)) return set_error(VFVM_ERR_SOLVER_INVALID_DAMPING, "adaptive: damp must be in (0, 1]");
    if (!opt->fixed && !(opt->rtol > 0.0 || opt->atol > 0.0))
        return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TOLERANCE, "adaptive: rtol or atol must be positive");
    vfvm_adaptive_stats as;
    std::memset(&as, 0, sizeof(as));

    // output times: sorted, without duplicates, ending at t_end (transient_adaptive.rs:103-110)
    const double t_end = opt->t_end;
    const double t_eps = 1e-12 * t_end;
    std::vector<double> outs;
    if (opt->num_out < 0) {
        for (int i = 0; i <= 10; ++i) outs.push_back(t_end * (double)i / 10.0);
    } else {
        if (opt->num_out > 0 && !opt->out_times) return set_error(VFVM_ERR_INVALID_ARGUMENT, "adaptive: out_times is NULL");
        outs.assign(opt->out_times, opt->out_times + opt->num_out);
    }
    outs.erase(std::remove_if(outs.begin(), outs.end(), [&](double t) { return !(t > t_eps && t <= t_end); }), outs.end());
    std::sort(outs.begin(), outs.end());
    {
        std::vector<double> d;
        for (double t : outs)
            if (d.empty() || std::fabs(t - d.back()) > t_eps) d.push_back(t);
        outs.swap(d);
    }
    if (!outs.empty() && outs.back() >= t_end - t_eps) outs.back() = t_end;
    else outs.push_back(t_end);

    // system, initial state (Dirichlet values at t = 0), previous state
    if (int st = vfvm_pattern_build(ctx, ops, nops, nullptr)) return st;
    if (c.flow && !c.fs.transient)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "adaptive: a flow list needs the velocity time operator (VEC_DOM_TIME) first");
    if (int st = call_update_hook(c, 0.0)) return st;
    if (int st = initial_dirichlet(c, ops, nops)) return st;
    System sy{c, ops, nops, c.n, c.flow};
    std::vector<ErrField> fields;
    auto excluded = [&](int32_t id) {
        for (int32_t k = 0; k < opt->num_exclude; ++k)
            if (opt->scl_exclude && opt->scl_exclude[k] == id) return true;
        return false;
    };
    // the error norm's fields (field_norms, scaling.rs:43-60), or with scale_off its blocks: every
    // scalar unknown and every velocity component on its own (_field_blocks, transient_adaptive.rs:355-370)
    if (c.flow) {
        if (c.scale_off)
            for (int k = 0; k < 3; ++k) fields.push_back({k * c.g.N, c.g.N, 0, -1});
        else
            fields.push_back({0, c.g.N, 1, c.fs.vel});
        if (!excluded(c.fs.pres)) fields.push_back({3 * c.g.N, c.g.N, 2, c.fs.pres});
        for (size_t k = 0; k < c.sblk.size(); ++k)
            if (!excluded(c.sblk[k].unk)) fields.push_back({(int64_t)(4 + k) * c.g.N, c.g.N, 0, c.sblk[k].unk});
    } else if (!excluded(c.unk)) {
        fields.push_back({0, c.g.N, 0, c.unk});
    }
    const int64_t n = sy.n;
    DevBuf<double> xn, xm, xi, xu, xd;
    for (DevBuf<double>* b : {&xn, &xm, &xi, &xu, &xd}) VFVM_HIP_TRY(b->alloc(n));
    if (int st = sy.pack(xn.p)) return st;
    // _update_func at the initial state, then _update_prev (transient_adaptive.rs:138-140)
    if (int st = call_update_hook(c, 0.0)) return st;
    if (int st = update_prev_all(c)) return st;

    int out_idx = 0;
    double t = 0.0;
    double dt = std::min(std::max(opt->dt_init, opt->dt_min), opt->dt_max);
    bool has_prev = false;
    double dt_prev = 0.0;
    double err_prev = 1.0;
    int n_reject = 0;
    bool no_grow = false;
    std::string last_solve_error;
    while (t < t_end - t_eps) {
        // land exactly on the next output time; avoid a tiny remainder step
        const double t_out = outs[out_idx];
        double dt_step = dt;
        bool hits = false;
        if (t + dt_step >= t_out - t_eps) {
            dt_step = t_out - t;
            hits = true;
        } else if (t + 1.5 * dt_step > t_out) {
            dt_step = 0.5 * (t_out - t);
        }
        const double t_new = hits ? t_out : t + dt_step;

        // Picard iterations at the end of the step from the last accepted state
        auto ta = clk::now();
        VFVM_HIP_TRY(hipMemcpyAsync(xi.p, xn.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        if (int st = call_update_hook(c, t_new)) return st;
        if (int st = sy.assemble(dt_step)) return st;
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        as.time_assembly += secs(ta, clk::now());
        bool conv = false;
        double res = NAN;
        for (int32_t iter = 0; iter < opt->max_iter; ++iter) {
            auto ti0 = clk::now();
            VFVM_HIP_TRY(hipMemcpyAsync(xu.p, xi.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
            vfvm_solve_stats ls;
            const int rs = gmres_solve(c, primal_view(c), *lin, xu.p, ls);
            as.linear_iterations += ls.iterations;
            auto ti1 = clk::now();
            as.time_solve += secs(ti0, ti1);
            if (rs != VFVM_OK) {
                // a failed linear solve rejects the step (transient_adaptive.rs:203-206)
                if (rs == VFVM_ERR_HIP) return rs;
                last_solve_error = vfvm_last_error();
                clear_error();
                break;
            }
            hipLaunchKernelGGL(k_damp_to, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, opt->damp, (const double*)xi.p,
                               (const double*)xu.p, xd.p);
            if (int st = sy.unpack(xd.p)) return st;
            if (int st = call_update_hook(c, t_new)) return st;
            if (int st = sy.assemble(dt_step)) return st;
            VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
            as.time_assembly += secs(ti1, clk::now());
            if (int st = sy.residual(xi.p, xd.p, &res)) return st;
            as.picard_iterations += 1;
            std::swap(xi, xd);
            if (!std::isfinite(res)) break;
            if (res < opt->tol) { conv = true; break; }
        }
        if (!conv) {
            // reject on Picard failure; unknowns back to the last accepted state
            as.rejected_picard += 1;
            n_reject += 1;
            if (opt->fixed || n_reject > opt->max_reject || dt_step <= opt->dt_min * (1.0 + 1e-12)) {
                if (stats) *stats = as;
                char buf[200];
                snprintf(buf, sizeof(buf), "adaptive: Picard iteration did not converge at step %d (residual %.3e)", as.accepted, res);
                return set_error(VFVM_ERR_TRANSIENT_DID_NOT_CONVERGE,
                                 std::string(buf) + (last_solve_error.empty() ? "" : "; last linear solve: " + last_solve_error));
            }
            if (int st = sy.unpack(xn.p)) return st;
            dt = std::max(dt_step * opt->fac_fail, opt->dt_min);
            no_grow = true;
            continue;
        }
        // error estimate
        double err = 0.0;
        if (!opt->fixed) {
            const double ratio = has_prev ? dt_step / dt_prev : 0.0;
            const double fac = has_prev ? dt_step / (2.0 * dt_step + dt_prev) : 0.5;
            if (int st = error_norm(sy, fields, xn.p, has_prev ? xm.p : nullptr, xi.p, ratio, fac, has_prev, opt->atol, opt->rtol, &err))
                return st;
        }
        if (err > 1.0 && dt_step > opt->dt_min * (1.0 + 1e-12)) {
            as.rejected_error += 1;
            n_reject += 1;
            if (n_reject > opt->max_reject) {
                if (stats) *stats = as;
                return set_error(VFVM_ERR_TRANSIENT_DID_NOT_CONVERGE, "adaptive: too many consecutive rejections");
            }
            if (int st = sy.unpack(xn.p)) return st;
            dt = std::max(dt_step * std::max(opt->safety * std::pow(err, -0.5), opt->fac_min), opt->dt_min);
            no_grow = true;
            continue;
        }
        // accept: the hook sees the new state with value_prev still at t - dt
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        if (hook) hook(user, t_new, dt_step);
        if (int st = update_prev_all(c)) return st;
        if (steps_out && as.accepted < max_steps) {
            steps_out[3 * as.accepted] = t_new;
            steps_out[3 * as.accepted + 1] = dt_step;
            steps_out[3 * as.accepted + 2] = err;
        }
        as.accepted += 1;
        std::swap(xm, xn);  // x_{n-1} := x_n
        std::swap(xn, xi);  // x_n := accepted state
        has_prev = true;
        dt_prev = dt_step;
        t = t_new;
        n_reject = 0;
        if (hits) {
            as.outputs += 1;
            out_idx += 1;
            if (out_idx >= (int)outs.size()) break;
        }
        // next step size (PI controller); a step shortened to hit an output time does not
        // limit the next one
        if (opt->fixed) { dt = opt->dt_init; continue; }
        const double e = std::max(err, 1e-10);
        double fac = opt->safety * std::pow(e, -0.35) * std::pow(std::max(err_prev, 1e-10), 0.2);
        fac = std::min(std::max(fac, opt->fac_min), opt->fac_max);
        if (no_grow) { fac = std::min(fac, 1.0); no_grow = false; }
        const double base = hits ? std::max(dt, dt_step) : dt_step;
        dt = std::min(std::max(base * fac, opt->dt_min), opt->dt_max);
        err_prev = e;
    }
    as.t_final = t;
    if (stats) *stats = as;
    return clear_error();
}
