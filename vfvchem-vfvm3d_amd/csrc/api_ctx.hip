// C ABI: device context, fields, operator pattern, assembly, solve and drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace vfvm;

// ------------------------------------------------------------------ errors
namespace {
thread_local std::string g_last_error;
}
int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
int clear_error() {
    g_last_error.clear();
    return VFVM_OK;
}
extern "C" const char* vfvm_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* vfvm_version(void) { return "vfvm-mi355x 0.1.0 (gfx950)"; }

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

__global__ void k_fill(int64_t n, double v, double* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = v;
}
// u[NP + nid[k]] = val[k]  (OpSclBndDirichlet::apply_initial, opscl_bnd_dir.rs:34-53)
// vec = f0 + df s per component (3N SoA), the affine closure of a vector function field
__global__ void k_affine(int64_t N, const double* __restrict__ f0, const double* __restrict__ df, const double* __restrict__ s,
                         double* __restrict__ vec) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= 3 * N) return;
    vec[i] = f0[i] + df[i] * s[i % N];
}
__global__ void k_scatter_bnd(int64_t cnt, int64_t NP, const int32_t* __restrict__ nid, const double* __restrict__ val,
                              double* __restrict__ u) {
    int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k < cnt) u[NP + nid[k]] = val[k];
}
// x_damp = (1 - d) x_iter + d x_udmp  (steady_base.rs:80)
__global__ void k_damp(int64_t n, double d, const double* __restrict__ xi, const double* __restrict__ xu, double* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = (1.0 - d) * xi[i] + d * xu[i];
}
// partials of sum (x_new - x_old)^2, max |x_old|, max |x_new|, non-finite count and
// sum x_old^2 (scaling.rs:113-135)
constexpr int PIC_PARTS = 5;
__global__ void k_picard(int64_t n, const double* __restrict__ xo, const double* __restrict__ xn, double* __restrict__ part) {
    __shared__ double s0[BLK], s1[BLK], s2[BLK], s3[BLK], s4[BLK];
    int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    double d2 = 0.0, mo = 0.0, mn = 0.0, nf = 0.0, o2 = 0.0;
    if (i < n) {
        double a = xo[i], b = xn[i];
        if (!isfinite(b)) nf = 1.0;
        double d = fabs(b - a);
        d2 = d * d;
        mo = fabs(a);
        mn = fabs(b);
        o2 = a * a;
    }
    s0[threadIdx.x] = d2; s1[threadIdx.x] = mo; s2[threadIdx.x] = mn; s3[threadIdx.x] = nf; s4[threadIdx.x] = o2;
    __syncthreads();
    for (int off = BLK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s0[threadIdx.x] += s0[threadIdx.x + off];
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + off]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + off]);
            s3[threadIdx.x] += s3[threadIdx.x + off];
            s4[threadIdx.x] += s4[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* p = part + PIC_PARTS * blockIdx.x;
        p[0] = s0[0]; p[1] = s1[0]; p[2] = s2[0]; p[3] = s3[0]; p[4] = s4[0];
    }
}

bool is_dom_op(int k) { return k >= 1 && k < 20; }
bool is_bnd_op(int k) { return k >= 20 && k < 40; }
bool is_itf_op(int k) { return k >= VFVM_OP_SCL_ITF_CONTINUITY && k <= VFVM_OP_SCL_ITF_ADVECTION_UNITY; }

struct Compiled {
    int nclass = 0;
    int32_t unk = -1;
    std::vector<std::vector<vfvm_op>> per_class;
    std::vector<uint8_t> has_edges, has_diag;
    std::vector<uint8_t> itf_cols;  // interface classes: ITF_* bits of the columns their rows hold
    bool has_time = false;
};

int check_scl(Ctx& c, int32_t id) { return id >= 0 && id < (int32_t)c.scl.size(); }

// validates an operator list and splits it by row class (points, then each boundary),
// keeping the list order inside each class (the reference's push order)
int compile_ops(Ctx& c, const vfvm_op* ops, int32_t nops, Compiled& out) {
    if (!c.has_mesh) return set_error(VFVM_ERR_INVALID_ARGUMENT, "no mesh uploaded");
    if (nops <= 0 || !ops) return set_error(VFVM_ERR_INVALID_ARGUMENT, "empty operator list");
    const DevGeom& g = c.g;
    out.nclass = g.cls_itf0 + 2 * g.num_itf;
    if (out.nclass > MAX_CLASSES) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "too many boundaries");
    out.per_class.assign(out.nclass, {});
    out.has_edges.assign(out.nclass, 0);
    out.has_diag.assign(out.nclass, 0);
    out.itf_cols.assign(out.nclass, 0);
    for (int32_t i = 0; i < nops; ++i) {
        const vfvm_op& op = ops[i];
        if (!check_scl(c, op.unk_id)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "operator " + std::to_string(i) + ": unknown id out of range");
        if (c.scl[op.unk_id].kind != VFVM_FIELD_UNKNOWN)
            return set_error(VFVM_ERR_OPERATOR_NON_UNKNOWN, "operator " + std::to_string(i) + ": row field is not an unknown (OperatorNonUnknownInLinearSystem)");
        if (out.unk < 0) out.unk = op.unk_id;
        else if (out.unk != op.unk_id) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "operators on more than one unknown are outside the built path");
        int cls;
        if (is_dom_op(op.kind)) {
            if (op.geom_id < 0 || op.geom_id >= g.num_dom) return set_error(VFVM_ERR_INVALID_ARGUMENT, "domain operator: domain id out of range");
            cls = op.geom_id;
        } else if (is_bnd_op(op.kind)) {
            if (op.geom_id < 0 || op.geom_id >= g.num_bnd) return set_error(VFVM_ERR_INVALID_ARGUMENT, "boundary operator: boundary id out of range");
            cls = g.num_dom + op.geom_id;
        } else if (is_itf_op(op.kind)) {
            if (op.geom_id < 0 || op.geom_id >= g.num_itf) return set_error(VFVM_ERR_INVALID_ARGUMENT, "interface operator: interface id out of range");
            cls = g.cls_itf0 + 2 * op.geom_id + (op.kind == VFVM_OP_SCL_ITF_CONTINUITY ? 0 : 1);
        } else {
            return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "unknown operator kind " + std::to_string(op.kind));
        }
        switch (op.kind) {
            case VFVM_OP_SCL_DOM_DIFFUSION:
            case VFVM_OP_SCL_BND_DIFFUSION:
                if (!check_scl(c, op.a)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "diffusion: coefficient id out of range");
                if (op.b != op.unk_id) {
                    if (check_scl(c, op.b) && c.scl[op.b].kind != VFVM_FIELD_UNKNOWN)
                        return set_error(VFVM_ERR_OPERATOR_NON_UNKNOWN, "diffusion: driving field is not an unknown (OperatorNonUnknownInLinearSystem)");
                    return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "diffusion: cross-diffusion into another unknown is outside the built path");
                }
                out.has_edges[cls] = 1;
                break;
            case VFVM_OP_SCL_DOM_ADVECTION_UNITY:
            case VFVM_OP_SCL_BND_ADVECTION_UNITY:
                if (op.a >= 0 && !check_scl(c, op.a)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "advection: diffusion id out of range");
                if (op.b < 0 || op.b >= (int32_t)c.flw.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "advection: flow id out of range");
                out.has_edges[cls] = 1;
                break;
            case VFVM_OP_SCL_DOM_ADVECTION:
            case VFVM_OP_SCL_BND_ADVECTION:
                if (!check_scl(c, op.a) || !check_scl(c, op.b)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "advection: weight or diffusion id out of range");
                if (op.c < 0 || op.c >= (int32_t)c.flw.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "advection: flow id out of range");
                out.has_edges[cls] = 1;
                break;
            case VFVM_OP_SCL_DOM_PRODUCTION:
            case VFVM_OP_SCL_BND_PRODUCTION:
            case VFVM_OP_SCL_DOM_TIME:
            case VFVM_OP_SCL_BND_TIME:
            case VFVM_OP_SCL_DOM_TIME_PREVIOUS:
            case VFVM_OP_SCL_BND_TIME_PREVIOUS:
                if (!check_scl(c, op.a)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "coefficient id out of range");
                out.has_diag[cls] = 1;
                if (op.kind != VFVM_OP_SCL_DOM_PRODUCTION && op.kind != VFVM_OP_SCL_BND_PRODUCTION) out.has_time = true;
                break;
            case VFVM_OP_SCL_DOM_TIME_UNITY:
            case VFVM_OP_SCL_BND_TIME_UNITY:
                out.has_diag[cls] = 1;
                out.has_time = true;
                break;
            case VFVM_OP_SCL_DOM_SOURCE:
            case VFVM_OP_SCL_BND_SOURCE:
                if (!check_scl(c, op.a)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "source id out of range");
                break;
            case VFVM_OP_SCL_BND_DIRICHLET:
            case VFVM_OP_SCL_BND_NEUMANN:
                if (op.a < 0 || op.a >= (int32_t)c.sbnd.size() || c.sbnd[op.a].bnd != op.geom_id)
                    return set_error(VFVM_ERR_INVALID_ARGUMENT, "boundary value id out of range or on another boundary");
                if (op.kind == VFVM_OP_SCL_BND_DIRICHLET) out.has_diag[cls] = 1;
                break;
            // interface operators (opscl_itf_*.rs): the columns their rows receive
            case VFVM_OP_SCL_ITF_CONTINUITY:
                out.itf_cols[cls] |= ITF_CA | ITF_CB;
                break;
            case VFVM_OP_SCL_ITF_DIFFUSION:
                if (!check_scl(c, op.a)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "interface diffusion: coefficient id out of range");
                if (op.b != op.unk_id) {
                    if (check_scl(c, op.b) && c.scl[op.b].kind != VFVM_FIELD_UNKNOWN)
                        return set_error(VFVM_ERR_OPERATOR_NON_UNKNOWN, "interface diffusion: driving field is not an unknown (OperatorNonUnknownInLinearSystem)");
                    return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "interface diffusion: cross-diffusion into another unknown is outside the built path");
                }
                out.itf_cols[cls] |= ITF_EA | ITF_EB | ITF_CA | ITF_CB;
                break;
            case VFVM_OP_SCL_ITF_PRODUCTION:
                if (!check_scl(c, op.a)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "interface production: coefficient id out of range");
                out.itf_cols[cls] |= ITF_CB;
                break;
            case VFVM_OP_SCL_ITF_SOURCE:
                if (!check_scl(c, op.a)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "interface source: id out of range");
                break;
            case VFVM_OP_SCL_ITF_TIME:
            case VFVM_OP_SCL_ITF_TIME_PREVIOUS:
                if (!check_scl(c, op.a)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "interface time: weight id out of range");
                out.itf_cols[cls] |= ITF_CA | ITF_CB;
                out.has_time = true;
                break;
            case VFVM_OP_SCL_ITF_TIME_UNITY:
                out.itf_cols[cls] |= ITF_CA | ITF_CB;
                out.has_time = true;
                break;
            case VFVM_OP_SCL_ITF_ADVECTION:
            case VFVM_OP_SCL_ITF_ADVECTION_UNITY: {
                const bool w = op.kind == VFVM_OP_SCL_ITF_ADVECTION;
                const int32_t fl = w ? op.c : op.b;
                if (w ? (!check_scl(c, op.a) || !check_scl(c, op.b)) : (op.a >= 0 && !check_scl(c, op.a)))
                    return set_error(VFVM_ERR_INVALID_ARGUMENT, "interface advection: weight or diffusion id out of range");
                if (fl < 0 || fl >= (int32_t)c.flw.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "interface advection: flow id out of range");
                if (op.value != 0.0 && op.value != 1.0)
                    return set_error(VFVM_ERR_INVALID_ARGUMENT, "interface advection: value must name the flowing side (0: a, 1: b)");
                out.itf_cols[cls] |= op.value == 0.0 ? (ITF_EA | ITF_CA) : (ITF_EB | ITF_CB);
            } break;
            default:
                return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "unknown operator kind " + std::to_string(op.kind));
        }
        if ((int)out.per_class[cls].size() >= MAX_OPS) return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "too many operators on one row class");
        out.per_class[cls].push_back(op);
    }
    return VFVM_OK;
}

int build_program(Ctx& c, const Compiled& cp, DevProgram& p) {
    std::memset(&p, 0, sizeof(p));
    SclDom& u = c.scl[cp.unk];
    for (int cls = 0; cls < cp.nclass; ++cls) {
        p.nops[cls] = (int32_t)cp.per_class[cls].size();
        for (size_t k = 0; k < cp.per_class[cls].size(); ++k) {
            const vfvm_op& op = cp.per_class[cls][k];
            DevOp& d = p.ops[cls][k];
            d.kind = op.kind;
            d.u_grad = u.grad.p;
            d.u_prev = u.value_prev.p;
            switch (op.kind) {
                case VFVM_OP_SCL_BND_DIRICHLET:
                case VFVM_OP_SCL_BND_NEUMANN:
                    d.f_val = c.sbnd[op.a].value.p;
                    break;
                case VFVM_OP_SCL_DOM_ADVECTION_UNITY:
                case VFVM_OP_SCL_BND_ADVECTION_UNITY:
                    d.has_diff = op.a >= 0;
                    if (op.a >= 0) { d.f_val = c.scl[op.a].value.p; d.f_grad = c.scl[op.a].grad.p; }
                    d.flux = c.flw[op.b].flux.p;
                    break;
                case VFVM_OP_SCL_DOM_ADVECTION:
                case VFVM_OP_SCL_BND_ADVECTION: {
                    const SclFlw& fl = c.flw[op.c];
                    d.has_diff = 1;
                    d.f_val = c.scl[op.b].value.p; d.f_grad = c.scl[op.b].grad.p;
                    d.w_val = c.scl[op.a].value.p; d.w_grad = c.scl[op.a].grad.p;
                    d.r_val = c.scl[fl.den].value.p; d.r_grad = c.scl[fl.den].grad.p;
                    d.flux = fl.flux.p;
                } break;
                case VFVM_OP_SCL_DOM_TIME_PREVIOUS:
                case VFVM_OP_SCL_BND_TIME_PREVIOUS:
                    d.f_val = c.scl[op.a].value_prev.p;
                    break;
                case VFVM_OP_SCL_DOM_TIME_UNITY:
                case VFVM_OP_SCL_BND_TIME_UNITY:
                case VFVM_OP_SCL_ITF_TIME_UNITY:
                case VFVM_OP_SCL_ITF_CONTINUITY:
                    break;
                case VFVM_OP_SCL_ITF_TIME_PREVIOUS:
                    d.f_val = c.scl[op.a].value_prev.p;
                    break;
                case VFVM_OP_SCL_ITF_ADVECTION: {
                    const SclFlw& fl = c.flw[op.c];
                    d.has_diff = 1;
                    d.f_val = c.scl[op.b].value.p; d.f_grad = c.scl[op.b].grad.p;
                    d.w_val = c.scl[op.a].value.p; d.w_grad = c.scl[op.a].grad.p;
                    d.r_val = c.scl[fl.den].value.p; d.r_grad = c.scl[fl.den].grad.p;
                    d.flux = fl.flux.p;
                    d.side = op.value == 0.0 ? 0 : 1;
                } break;
                case VFVM_OP_SCL_ITF_ADVECTION_UNITY:
                    d.has_diff = op.a >= 0;
                    if (op.a >= 0) { d.f_val = c.scl[op.a].value.p; d.f_grad = c.scl[op.a].grad.p; }
                    d.flux = c.flw[op.b].flux.p;
                    d.side = op.value == 0.0 ? 0 : 1;
                    break;
                default:
                    d.f_val = c.scl[op.a].value.p;
                    d.f_grad = c.scl[op.a].grad.p;
            }
        }
    }
    return VFVM_OK;
}

}  // namespace

int vfvm::sync_err(Ctx& c) {
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    return VFVM_OK;
}

// ------------------------------------------------------------------ context
extern "C" int vfvm_create(int32_t device, vfvm_ctx** out) {
    if (!out) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_create: out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return set_error(VFVM_ERR_HIP, "vfvm_create: no HIP device available");
    if (device < 0 || device >= count) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_create: device index out of range");
    VFVM_HIP_TRY(hipSetDevice(device));
    auto* h = new vfvm_ctx();
    h->c.device = device;
    const char* gr = getenv("VFVM_GRAPHS");
    h->c.graphs_enabled = !(gr && gr[0] == '0');
    e = hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete h; return set_error(VFVM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    // pinned readback buffer of the solver's control word and norms
    e = hipHostMalloc((void**)&h->c.h_pinned, 4096, hipHostMallocDefault);
    if (e != hipSuccess) { vfvm_destroy(h); return set_error(VFVM_ERR_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(e)); }
    // device address of this GPU's no-allocation scratch array (once per device)
    if (int st = noalloc_setup(device)) { vfvm_destroy(h); return st; }
    *out = h;
    return clear_error();
}

extern "C" void vfvm_destroy(vfvm_ctx* ctx) {
    if (!ctx) return;
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (auto e : c.ev_pool) (void)hipEventDestroy(e);
    for (auto e : c.ctrl_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& g : c.graphs)
        for (auto e : g.chunks) (void)hipGraphExecDestroy(e);
    if (c.h_pinned) (void)hipHostFree(c.h_pinned);
    hipStream_t s = c.stream;
    delete ctx;  // frees device buffers
    if (s) (void)hipStreamDestroy(s);
}

extern "C" int vfvm_mesh_upload(vfvm_ctx* ctx, const vfvm_mesh* mesh) {
    CTX_GUARD(ctx);
    const Mesh* mp = vfvm_mesh_ptr(mesh);
    if (!mp || !mp->has_dual) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_upload: mesh without dual");
    const Mesh& m = *mp;
    // reset everything that depends on the mesh
    c.scl.clear(); c.vec.clear(); c.sbnd.clear(); c.flw.clear(); c.vbndv.clear();
    c.scale_scl.clear(); c.scale_vec.clear(); c.vaff.clear();
    c.lin.active = false;
    c.has_pattern = false; c.has_matrix = false; c.ilu.ready = false; c.flow = false;
    c.has_transpose = false; c.tilu.ready = false; c.tape.valid = false;
    DevGeom& g = c.g;
    g.rev.reset();
    g.gct_ptr.reset(); g.gct_slot.reset(); g.gct_own.reset();
    g.N = m.num_tot(); g.NP = m.num_pts; g.E = m.num_edges(); g.S = m.num_subfaces();
    g.num_bnd = (int32_t)m.bnd_name.size();
    hipStream_t s = c.stream;
    VFVM_HIP_TRY(g.row_ptr.upload(m.row_ptr.data(), m.row_ptr.size(), s));
    VFVM_HIP_TRY(g.ngh.upload(m.ngh.data(), m.ngh.size(), s));
    std::vector<int32_t> erow(g.E);
    for (int64_t v = 0; v < g.N; ++v)
        for (int32_t e = m.row_ptr[v]; e < m.row_ptr[v + 1]; ++e) erow[e] = (int32_t)v;
    VFVM_HIP_TRY(g.erow.upload(erow.data(), erow.size(), s));
    VFVM_HIP_TRY(g.dist.upload(m.dist.data(), m.dist.size(), s));
    VFVM_HIP_TRY(g.area.upload(m.area.data(), m.area.size(), s));
    VFVM_HIP_TRY(g.norm.upload(m.norm.data(), m.norm.size(), s));
    VFVM_HIP_TRY(g.skew.upload(m.skew.data(), m.skew.size(), s));
    VFVM_HIP_TRY(g.gw.upload(m.gw.data(), m.gw.size(), s));
    VFVM_HIP_TRY(g.vol.upload(m.vol.data(), m.vol.size(), s));
    VFVM_HIP_TRY(g.cskew.upload(m.cskew.data(), m.cskew.size(), s));
    VFVM_HIP_TRY(g.bsf_ptr.upload(m.bsf_ptr.data(), m.bsf_ptr.size(), s));
    VFVM_HIP_TRY(g.bsf_vj.upload(m.bsf_vj.data(), m.bsf_vj.size(), s));
    VFVM_HIP_TRY(g.bsf_vk.upload(m.bsf_vk.data(), m.bsf_vk.size(), s));
    VFVM_HIP_TRY(g.bsf_area.upload(m.bsf_area.data(), m.bsf_area.size(), s));
    VFVM_HIP_TRY(g.bsf_wgt.upload(m.bsf_wgt.data(), m.bsf_wgt.size(), s));
    VFVM_HIP_TRY(g.bsf_norm.upload(m.bsf_norm.data(), m.bsf_norm.size(), s));
    g.G = m.has_grad_coeff ? (int64_t)m.gc_vid.size() : 0;
    if (m.has_grad_coeff) {
        VFVM_HIP_TRY(g.gc_ptr.upload(m.gc_ptr.data(), m.gc_ptr.size(), s));
        VFVM_HIP_TRY(g.gc_vid.upload(m.gc_vid.data(), m.gc_vid.size(), s));
        VFVM_HIP_TRY(g.gc_c.upload(m.gc_c.data(), m.gc_c.size(), s));
        g.h_gc_ptr = m.gc_ptr;
        g.h_gc_vid = m.gc_vid;
    } else {
        g.gc_ptr.reset(); g.gc_vid.reset(); g.gc_c.reset();
        g.h_gc_ptr.clear(); g.h_gc_vid.clear();
    }
    std::vector<int32_t> vbnd(g.N, -1), bidx(g.N, -1);
    for (int32_t b = 0; b < g.num_bnd; ++b)
        for (int64_t k = m.bnd_ptr[b]; k < m.bnd_ptr[b + 1]; ++k) {
            int64_t v = m.num_pts + m.bnd_nid[k];
            vbnd[v] = b;
            bidx[v] = (int32_t)(k - m.bnd_ptr[b]);
        }
    VFVM_HIP_TRY(g.vbnd.upload(vbnd.data(), vbnd.size(), s));
    VFVM_HIP_TRY(g.bidx.upload(bidx.data(), bidx.size(), s));
    // operator classes: points of domain d -> d, boundary b -> num_dom + b, interface k's
    // side-a / side-b copies -> cls_itf0 + 2k / + 2k + 1
    g.num_dom = m.num_dom();
    g.num_itf = m.num_itf();
    g.multi = g.num_dom > 1 || g.num_itf > 0;
    g.cls_itf0 = g.num_dom + g.num_bnd;
    if (g.cls_itf0 + 2 * g.num_itf > MAX_CLASSES)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "vfvm_mesh_upload: more than " + std::to_string(MAX_CLASSES) +
                                                        " domains + boundaries + 2 x interfaces");
    // (a single-domain mesh treats a node on no boundary as a point, as the class of its rows
    // always did)
    g.h_vcls.assign(g.N, g.multi ? -1 : 0);
    for (int32_t d = 0; d < g.num_dom; ++d) {
        const int64_t v0 = m.dom_ptr.empty() ? 0 : m.dom_ptr[d];
        const int64_t np = m.dom_npts.empty() ? m.num_pts : m.dom_npts[d];
        for (int64_t v = v0; v < v0 + np; ++v) g.h_vcls[v] = d;
    }
    for (int64_t v = 0; v < g.N; ++v)
        if (vbnd[v] >= 0) g.h_vcls[v] = g.num_dom + vbnd[v];
    g.K = (int64_t)m.itf_va.size();
    g.h_itf_va = m.itf_va;
    g.h_itf_vb = m.itf_vb;
    g.h_itf_of.assign(g.K, 0);
    for (int32_t k = 0; k < g.num_itf; ++k)
        for (int64_t q = m.itf_ptr[k]; q < m.itf_ptr[k + 1]; ++q) {
            g.h_itf_of[q] = k;
            g.h_vcls[m.itf_va[q]] = g.cls_itf0 + 2 * k;
            g.h_vcls[m.itf_vb[q]] = g.cls_itf0 + 2 * k + 1;
        }
    for (int64_t v = 0; v < g.N; ++v)
        if (g.h_vcls[v] < 0) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_mesh_upload: node " + std::to_string(v) + " is on no boundary or interface");
    VFVM_HIP_TRY(g.vcls.upload(g.h_vcls.data(), g.h_vcls.size(), s));
    VFVM_HIP_TRY(g.itf_va.upload(g.h_itf_va.data(), g.h_itf_va.size(), s));
    VFVM_HIP_TRY(g.itf_vb.upload(g.h_itf_vb.data(), g.h_itf_vb.size(), s));
    VFVM_HIP_TRY(g.itf_of.upload(g.h_itf_of.data(), g.h_itf_of.size(), s));
    g.bnd_ptr = m.bnd_ptr;
    g.bnd_nid = m.bnd_nid;
    g.h_row_ptr = m.row_ptr;
    g.h_ngh = m.ngh;
    g.h_pos = m.pos;
    if (int st = sync_err(c)) return st;
    c.has_mesh = true;
    return clear_error();
}

// ------------------------------------------------------------------ fields
extern "C" int vfvm_scldom_add(vfvm_ctx* ctx, int32_t kind, double cnst, const double* values, int32_t* id) {
    CTX_GUARD(ctx);
    if (!c.has_mesh || !id || kind < 0 || kind > 2) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_scldom_add: invalid argument");
    const int64_t N = c.g.N;
    c.scl.emplace_back();
    SclDom& f = c.scl.back();
    f.kind = kind;
    f.cnst = cnst;
    VFVM_HIP_TRY(f.value.alloc(N));
    VFVM_HIP_TRY(f.value_prev.alloc(N));
    VFVM_HIP_TRY(f.grad.alloc(3 * N));
    VFVM_HIP_TRY(hipMemsetAsync(f.grad.p, 0, 3 * N * sizeof(double), c.stream));
    if (values) VFVM_HIP_TRY(hipMemcpyAsync(f.value.p, values, N * sizeof(double), hipMemcpyHostToDevice, c.stream));
    else hipLaunchKernelGGL(k_fill, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, cnst, f.value.p);
    VFVM_HIP_TRY(hipMemcpyAsync(f.value_prev.p, f.value.p, N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    if (int st = sync_err(c)) return st;
    *id = (int32_t)c.scl.size() - 1;
    return clear_error();
}

extern "C" int vfvm_scldom_set(vfvm_ctx* ctx, int32_t id, const double* values) {
    CTX_GUARD(ctx);
    if (!check_scl(c, id) || !values) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_scldom_set: invalid argument");
    VFVM_HIP_TRY(hipMemcpyAsync(c.scl[id].value.p, values, c.g.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
    c.scl[id].grad_valid = false;
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_scldom_get(vfvm_ctx* ctx, int32_t id, double* values) {
    CTX_GUARD(ctx);
    if (!check_scl(c, id) || !values) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_scldom_get: invalid argument");
    VFVM_HIP_TRY(hipMemcpyAsync(values, c.scl[id].value.p, c.g.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_scldom_update_prev(vfvm_ctx* ctx, int32_t id) {
    CTX_GUARD(ctx);
    if (!check_scl(c, id)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_scldom_update_prev: id out of range");
    VFVM_HIP_TRY(hipMemcpyAsync(c.scl[id].value_prev.p, c.scl[id].value.p, c.g.N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_vecdom_add(vfvm_ctx* ctx, int32_t kind, const double* cnst3, const double* values, int32_t* id) {
    CTX_GUARD(ctx);
    if (!c.has_mesh || !id || kind < 0 || kind > 2 || (!values && !cnst3))
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_vecdom_add: invalid argument");
    const int64_t N = c.g.N;
    c.vec.emplace_back();
    VecDom& f = c.vec.back();
    f.kind = kind;
    VFVM_HIP_TRY(f.value.alloc(3 * N));
    VFVM_HIP_TRY(f.grad.alloc(9 * N));
    VFVM_HIP_TRY(hipMemsetAsync(f.grad.p, 0, 9 * N * sizeof(double), c.stream));
    if (values) VFVM_HIP_TRY(hipMemcpyAsync(f.value.p, values, 3 * N * sizeof(double), hipMemcpyHostToDevice, c.stream));
    else
        for (int k = 0; k < 3; ++k) hipLaunchKernelGGL(k_fill, dim3(grid_for(N)), dim3(BLK), 0, c.stream, N, cnst3[k], f.value.p + k * N);
    VFVM_HIP_TRY(f.value_prev.alloc(3 * N));
    VFVM_HIP_TRY(hipMemcpyAsync(f.value_prev.p, f.value.p, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    if (int st = sync_err(c)) return st;
    *id = (int32_t)c.vec.size() - 1;
    return clear_error();
}

extern "C" int vfvm_vecdom_set(vfvm_ctx* ctx, int32_t id, const double* values) {
    CTX_GUARD(ctx);
    if (id < 0 || id >= (int32_t)c.vec.size() || !values) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_vecdom_set: invalid argument");
    VFVM_HIP_TRY(hipMemcpyAsync(c.vec[id].value.p, values, 3 * c.g.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
    c.vec[id].grad_valid = false;
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_sclbnd_add(vfvm_ctx* ctx, int32_t bnd_id, double cnst, const double* values, int32_t* id) {
    CTX_GUARD(ctx);
    if (!c.has_mesh || !id || bnd_id < 0 || bnd_id >= c.g.num_bnd) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_sclbnd_add: invalid argument");
    const int64_t cnt = c.g.bnd_ptr[bnd_id + 1] - c.g.bnd_ptr[bnd_id];
    c.sbnd.emplace_back();
    SclBnd& f = c.sbnd.back();
    f.bnd = bnd_id;
    VFVM_HIP_TRY(f.value.alloc(std::max<int64_t>(cnt, 1)));
    if (values) VFVM_HIP_TRY(hipMemcpyAsync(f.value.p, values, cnt * sizeof(double), hipMemcpyHostToDevice, c.stream));
    else hipLaunchKernelGGL(k_fill, dim3(grid_for(std::max<int64_t>(cnt, 1))), dim3(BLK), 0, c.stream, std::max<int64_t>(cnt, 1), cnst, f.value.p);
    if (int st = sync_err(c)) return st;
    *id = (int32_t)c.sbnd.size() - 1;
    return clear_error();
}

extern "C" int vfvm_sclbnd_set(vfvm_ctx* ctx, int32_t id, const double* values) {
    CTX_GUARD(ctx);
    if (id < 0 || id >= (int32_t)c.sbnd.size() || !values) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_sclbnd_set: invalid argument");
    const int32_t b = c.sbnd[id].bnd;
    const int64_t cnt = c.g.bnd_ptr[b + 1] - c.g.bnd_ptr[b];
    VFVM_HIP_TRY(hipMemcpyAsync(c.sbnd[id].value.p, values, cnt * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_sclflw_add(vfvm_ctx* ctx, int32_t vel_id, int32_t pres_id, int32_t den_id, int32_t visc_id, int32_t fce_id, int32_t* id) {
    CTX_GUARD(ctx);
    if (!id || vel_id < 0 || vel_id >= (int32_t)c.vec.size() || !check_scl(c, pres_id) || !check_scl(c, den_id) ||
        !check_scl(c, visc_id) || fce_id < 0 || fce_id >= (int32_t)c.vec.size())
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_sclflw_add: field id out of range");
    // a constant pressure: prescribed flow (face flux from the velocity alone); an unknown
    // pressure: a flow solved by the flow operators (Rhie-Chow face flux)
    if (c.scl[pres_id].kind == VFVM_FIELD_FUNCTION)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "vfvm_sclflw_add: the pressure must be a constant (prescribed flow) or an unknown (solved flow)");
    if (c.scl[pres_id].kind == VFVM_FIELD_UNKNOWN && c.vec[vel_id].kind != VFVM_FIELD_UNKNOWN)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "vfvm_sclflw_add: a solved flow needs an unknown velocity");
    c.flw.emplace_back();
    SclFlw& f = c.flw.back();
    f.vel = vel_id; f.pres = pres_id; f.den = den_id; f.visc = visc_id; f.fce = fce_id;
    VFVM_HIP_TRY(f.flux.alloc(c.g.E));
    *id = (int32_t)c.flw.size() - 1;
    return clear_error();
}

extern "C" int vfvm_vecdom_get(vfvm_ctx* ctx, int32_t id, double* values) {
    CTX_GUARD(ctx);
    if (id < 0 || id >= (int32_t)c.vec.size() || !values) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_vecdom_get: invalid argument");
    VFVM_HIP_TRY(hipMemcpyAsync(values, c.vec[id].value.p, 3 * c.g.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_update_flow(vfvm_ctx* ctx, int32_t flw_id) {
    CTX_GUARD(ctx);
    if (int st = flow_update(c, flw_id)) return st;
    return clear_error();
}

extern "C" int vfvm_sclflw_get(vfvm_ctx* ctx, int32_t flw_id, int32_t what, double* out) {
    CTX_GUARD(ctx);
    if (flw_id < 0 || flw_id >= (int32_t)c.flw.size() || !out || what < 0 || what > 2)
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_sclflw_get: invalid argument");
    const SclFlw& f = c.flw[flw_id];
    const DevBuf<double>& src = what == 0 ? f.flux : (what == 1 ? f.rc : f.mom);
    const size_t cnt = what == 0 ? (size_t)c.g.E : (what == 1 ? (size_t)c.g.N : 3 * (size_t)c.g.N);
    if (src.n < cnt || !src.p) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_sclflw_get: the flow data has not been computed");
    VFVM_HIP_TRY(hipMemcpyAsync(out, src.p, cnt * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_vecbnd_add(vfvm_ctx* ctx, int32_t bnd_id, const double* cnst3, const double* values, int32_t* id) {
    CTX_GUARD(ctx);
    if (!c.has_mesh || !id || bnd_id < 0 || bnd_id >= c.g.num_bnd || (!values && !cnst3))
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_vecbnd_add: invalid argument");
    const int64_t cnt = c.g.bnd_ptr[bnd_id + 1] - c.g.bnd_ptr[bnd_id];
    c.vbndv.emplace_back();
    VecBnd& f = c.vbndv.back();
    f.bnd = bnd_id;
    const int64_t m = std::max<int64_t>(cnt, 1);
    VFVM_HIP_TRY(f.value.alloc(3 * m));
    if (values) VFVM_HIP_TRY(hipMemcpyAsync(f.value.p, values, 3 * cnt * sizeof(double), hipMemcpyHostToDevice, c.stream));
    else
        for (int k = 0; k < 3; ++k) hipLaunchKernelGGL(k_fill, dim3(grid_for(m)), dim3(BLK), 0, c.stream, m, cnst3[k], f.value.p + k * m);
    if (int st = sync_err(c)) return st;
    *id = (int32_t)c.vbndv.size() - 1;
    return clear_error();
}

extern "C" int vfvm_vecbnd_set(vfvm_ctx* ctx, int32_t id, const double* values) {
    CTX_GUARD(ctx);
    if (id < 0 || id >= (int32_t)c.vbndv.size() || !values) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_vecbnd_set: invalid argument");
    const int32_t b = c.vbndv[id].bnd;
    const int64_t cnt = c.g.bnd_ptr[b + 1] - c.g.bnd_ptr[b];
    if (cnt > 0) VFVM_HIP_TRY(hipMemcpyAsync(c.vbndv[id].value.p, values, 3 * cnt * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

// ------------------------------------------------------------------ pattern (LinearBase::update_sparse)
extern "C" int vfvm_pattern_build(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, int64_t* nnz) {
    CTX_GUARD(ctx);
    c.lin.active = false;
    c.has_pattern = false;
    c.has_matrix = false;
    c.ilu.ready = false;
    c.has_transpose = false;
    c.tilu.ready = false;
    c.tape.valid = false;
    c.flow = false;
    if (is_flow_list(ops, nops)) {
        // coupled velocity-pressure system (SteadyFlow)
        if (int st = flow_pattern_build(c, ops, nops)) return st;
        c.has_pattern = true;
        if (nnz) *nnz = c.nnz;
        return clear_error();
    }
    Compiled cp;
    if (int st = compile_ops(c, ops, nops, cp)) return st;
    const DevGeom& g = c.g;
    const int64_t N = g.N;
    // single unknown: block start 0 (SteadyHeat::initial_start, steady_heat.rs:68-78)
    c.unk = cp.unk;
    c.n = N;
    c.h_mrow.assign(N + 1, 0);
    c.h_mcol.clear();
    std::vector<int32_t> eslot(g.E, -1), dslot(N, -1);
    // interface node of each copy (-1: none)
    std::vector<int32_t> itf_q(g.multi ? N : 0, -1);
    for (int64_t q = 0; q < g.K; ++q) itf_q[g.h_itf_va[q]] = itf_q[g.h_itf_vb[q]] = (int32_t)q;
    std::vector<int32_t> islot(4 * g.K, -1);
    std::vector<int32_t> cols;
    for (int64_t v = 0; v < N; ++v) {
        const int cls = g.h_vcls[v];
        const int32_t es = g.h_row_ptr[v], ee = g.h_row_ptr[v + 1];
        if (cls >= g.cls_itf0) {
            // interface rows: columns of the operators on this side's class, ascending
            // (opscl_itf_cont.rs:50-53 for side a; opscl_itf_diff.rs:73-108 and the other
            // operators load into side b's rows)
            const int32_t q = itf_q[v];
            const int32_t va = g.h_itf_va[q], vb = g.h_itf_vb[q];
            const uint8_t m = cp.itf_cols[cls];
            const bool side_b = ((cls - g.cls_itf0) & 1) != 0;
            cols.clear();
            if (side_b) {
                if (m & ITF_EA)
                    for (int32_t e = g.h_row_ptr[va]; e < g.h_row_ptr[va + 1]; ++e) cols.push_back(g.h_ngh[e]);
                if (m & ITF_EB)
                    for (int32_t e = g.h_row_ptr[vb]; e < g.h_row_ptr[vb + 1]; ++e) cols.push_back(g.h_ngh[e]);
            }
            if (m & ITF_CA) cols.push_back(va);
            if (m & ITF_CB) cols.push_back(vb);
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            const int32_t base = (int32_t)c.h_mcol.size();
            c.h_mcol.insert(c.h_mcol.end(), cols.begin(), cols.end());
            auto slot = [&](int32_t col) -> int32_t {
                auto it = std::lower_bound(cols.begin(), cols.end(), col);
                return (it != cols.end() && *it == col) ? base + (int32_t)(it - cols.begin()) : -1;
            };
            if (side_b) {
                islot[4 * q + 2] = slot(va);
                islot[4 * q + 3] = slot(vb);
                if (m & ITF_EA)
                    for (int32_t e = g.h_row_ptr[va]; e < g.h_row_ptr[va + 1]; ++e) eslot[e] = slot(g.h_ngh[e]);
                if (m & ITF_EB)
                    for (int32_t e = g.h_row_ptr[vb]; e < g.h_row_ptr[vb + 1]; ++e) eslot[e] = slot(g.h_ngh[e]);
            } else {
                islot[4 * q + 0] = slot(va);
                islot[4 * q + 1] = slot(vb);
            }
        } else if (cp.has_edges[cls]) {
            // {v} u ngh(v), ascending (opscl_dom_diff.rs:84-85)
            bool placed = false;
            for (int32_t e = es; e < ee; ++e) {
                const int32_t n = g.h_ngh[e];
                if (!placed && n > v) { dslot[v] = (int32_t)c.h_mcol.size(); c.h_mcol.push_back((int32_t)v); placed = true; }
                eslot[e] = (int32_t)c.h_mcol.size();
                c.h_mcol.push_back(n);
            }
            if (!placed) { dslot[v] = (int32_t)c.h_mcol.size(); c.h_mcol.push_back((int32_t)v); }
        } else if (cp.has_diag[cls]) {
            // {v}: Dirichlet, production- or time-only rows (opscl_bnd_dir.rs:73-74)
            dslot[v] = (int32_t)c.h_mcol.size();
            c.h_mcol.push_back((int32_t)v);
        }
        if ((int64_t)c.h_mcol.size() > INT32_MAX) return set_error(VFVM_ERR_INVALID_ARGUMENT, "pattern exceeds 32-bit indices");
        c.h_mrow[v + 1] = (int32_t)c.h_mcol.size();
    }
    c.nnz = (int64_t)c.h_mcol.size();
    hipStream_t s = c.stream;
    VFVM_HIP_TRY(c.mrow.upload(c.h_mrow.data(), c.h_mrow.size(), s));
    VFVM_HIP_TRY(c.mcol.upload(c.h_mcol.data(), c.h_mcol.size(), s));
    VFVM_HIP_TRY(c.eslot.upload(eslot.data(), eslot.size(), s));
    VFVM_HIP_TRY(c.dslot.upload(dslot.data(), dslot.size(), s));
    VFVM_HIP_TRY(c.islot.upload(islot.data(), islot.size(), s));
    VFVM_HIP_TRY(c.vals.alloc(std::max<int64_t>(c.nnz, 1)));
    VFVM_HIP_TRY(c.b.alloc(N));
    DevProgram p;
    build_program(c, cp, p);
    VFVM_HIP_TRY(c.prog.upload(&p, 1, s));
    if (int st = sync_err(c)) return st;
    c.pattern_ops.assign(ops, ops + nops);
    c.has_pattern = true;
    c.has_matrix = false;
    c.ilu.ready = false;
    c.has_transpose = false;
    c.tilu.ready = false;
    c.tape.valid = false;
    if (nnz) *nnz = c.nnz;
    return clear_error();
}

extern "C" int vfvm_system_info(vfvm_ctx* ctx, int64_t* n, int64_t* nnz) {
    CTX_GUARD(ctx);
    if (!c.has_pattern) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_system_info: no pattern built");
    if (n) *n = c.n;
    if (nnz) *nnz = c.nnz;
    return clear_error();
}

extern "C" int vfvm_pattern_get(vfvm_ctx* ctx, int32_t* row_ptr, int32_t* col) {
    CTX_GUARD(ctx);
    if (!c.has_pattern) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_pattern_get: no pattern built");
    if (row_ptr) std::memcpy(row_ptr, c.h_mrow.data(), c.h_mrow.size() * sizeof(int32_t));
    if (col) std::memcpy(col, c.h_mcol.data(), c.h_mcol.size() * sizeof(int32_t));
    return clear_error();
}

// ------------------------------------------------------------------ assembly
// Gradients of every non-constant field a scalar operator list reads (ScalarDomain::update_grad)
// and the face mass flux of the flows it uses; skip_flw: a flow whose data is already current.
int vfvm::prepare_scalar_fields(Ctx& c, const vfvm_op* ops, int32_t nops, int32_t skip_flw) {
    // gradients of every non-constant field the operators read (ScalarDomain::update_grad)
    std::vector<uint8_t> need_s(c.scl.size(), 0), need_v(c.vec.size(), 0), need_f(c.flw.size(), 0);
    for (int32_t i = 0; i < nops; ++i) {
        const vfvm_op& op = ops[i];
        switch (op.kind) {
            case VFVM_OP_SCL_DOM_PRODUCTION:
            case VFVM_OP_SCL_BND_PRODUCTION:
                need_s[op.a] = 1; need_s[op.unk_id] = 1; break;
            case VFVM_OP_SCL_DOM_ADVECTION_UNITY:
            case VFVM_OP_SCL_BND_ADVECTION_UNITY:
                if (op.a >= 0) need_s[op.a] = 1;
                need_f[op.b] = 1; break;
            case VFVM_OP_SCL_DOM_ADVECTION:
            case VFVM_OP_SCL_BND_ADVECTION:
            case VFVM_OP_SCL_ITF_ADVECTION:
                need_s[op.a] = 1; need_s[op.b] = 1; need_f[op.c] = 1; break;
            case VFVM_OP_SCL_ITF_PRODUCTION:
                need_s[op.a] = 1; need_s[op.unk_id] = 1; break;
            case VFVM_OP_SCL_ITF_ADVECTION_UNITY:
                if (op.a >= 0) need_s[op.a] = 1;
                need_f[op.b] = 1; break;
            case VFVM_OP_SCL_ITF_CONTINUITY:
            case VFVM_OP_SCL_ITF_TIME_PREVIOUS:
            case VFVM_OP_SCL_ITF_TIME_UNITY:
            case VFVM_OP_SCL_BND_DIRICHLET:
            case VFVM_OP_SCL_BND_NEUMANN:
            case VFVM_OP_SCL_DOM_TIME_PREVIOUS:
            case VFVM_OP_SCL_BND_TIME_PREVIOUS:
            case VFVM_OP_SCL_DOM_TIME_UNITY:
            case VFVM_OP_SCL_BND_TIME_UNITY:
                break;
            default:
                need_s[op.a] = 1;
        }
    }
    // a flow's flux needs its velocity and density gradients; the flow whose data is
    // current (skip_flw) had them computed with it
    for (size_t f = 0; f < c.flw.size(); ++f)
        if (need_f[f] && (int32_t)f != skip_flw) { need_s[c.flw[f].den] = 1; need_v[c.flw[f].vel] = 1; }
    // a function field's gradient is kept until its values are set again; an unknown's
    // values change with every solve
    for (size_t i = 0; i < c.scl.size(); ++i) {
        SclDom& f = c.scl[i];
        if (!need_s[i] || f.kind == VFVM_FIELD_CONSTANT || (f.kind == VFVM_FIELD_FUNCTION && f.grad_valid)) continue;
        if (int st = launch_grad(c, f.value.p, f.grad.p)) return st;
        f.grad_valid = f.kind == VFVM_FIELD_FUNCTION;
    }
    for (size_t i = 0; i < c.vec.size(); ++i) {
        VecDom& f = c.vec[i];
        if (!need_v[i] || f.kind == VFVM_FIELD_CONSTANT || (f.kind == VFVM_FIELD_FUNCTION && f.grad_valid)) continue;
        if (int st = launch_grad_vec(c, f.value.p, f.grad.p)) return st;
        f.grad_valid = f.kind == VFVM_FIELD_FUNCTION;
    }
    // face mass flux: a prescribed flow's from its velocity, a solved flow's (unknown
    // pressure) from the Rhie-Chow interpolation of its current fields
    for (size_t f = 0; f < c.flw.size(); ++f) {
        if (!need_f[f] || (int32_t)f == skip_flw) continue;
        if (c.scl[c.flw[f].pres].kind == VFVM_FIELD_UNKNOWN) {
            if (int st = flow_update(c, (int32_t)f)) return st;
        } else if (int st = launch_flux_prescribed(c, c.flw[f])) {
            return st;
        }
    }
    return VFVM_OK;
}

int vfvm::assemble_impl(Ctx& c, const vfvm_op* ops, int32_t nops, double dt) {
    if (!c.has_pattern) return set_error(VFVM_ERR_INVALID_ARGUMENT, "assemble: no pattern built (call vfvm_pattern_build)");
    if (c.lin.active)
        return set_error(VFVM_ERR_PATTERN_MISMATCH, "assemble: the system was built from triplets (vfvm_linear_update_sparse)");
    if (c.flow || is_flow_list(ops, nops)) return flow_assemble_dt(c, ops, nops, dt);
    Compiled cp;
    if (int st = compile_ops(c, ops, nops, cp)) return st;
    if (cp.unk != c.unk) return set_error(VFVM_ERR_PATTERN_MISMATCH, "assemble: operators act on another unknown than the pattern");
    {
        Compiled pp;
        if (int st = compile_ops(c, c.pattern_ops.data(), (int32_t)c.pattern_ops.size(), pp)) return st;
        if (pp.has_edges != cp.has_edges || pp.has_diag != cp.has_diag || pp.itf_cols != cp.itf_cols)
            return set_error(VFVM_ERR_PATTERN_MISMATCH, "assemble: operators do not fit the built pattern");
    }
    if (cp.has_time && !(dt > 0.0)) return set_error(VFVM_ERR_OPERATOR_USE_APPLY_TIME, "time operators need dt > 0 (OperatorUseApplyTime)");
    if (int st = prepare_scalar_fields(c, ops, nops, -1)) return st;
    DevProgram p;
    build_program(c, cp, p);
    VFVM_HIP_TRY(hipMemcpyAsync(c.prog.p, &p, sizeof(p), hipMemcpyHostToDevice, c.stream));
    if (int st = launch_assemble(c, dt)) return st;
    c.has_matrix = true;
    return VFVM_OK;
}

int vfvm::scalar_classes(Ctx& c, const vfvm_op* ops, int32_t nops, int32_t* unk, std::vector<uint8_t>& has_edges,
                         std::vector<uint8_t>& has_diag, bool* has_time) {
    Compiled cp;
    if (int st = compile_ops(c, ops, nops, cp)) return st;
    *unk = cp.unk;
    has_edges = cp.has_edges;
    has_diag = cp.has_diag;
    *has_time = cp.has_time;
    return VFVM_OK;
}

// One scalar unknown's rows inside a coupled system: the scalar assembly kernel writes the
// block's values through its own edge/diagonal slot maps and its b segment.
int vfvm::scalar_block_assemble(Ctx& c, ScalarBlock& blk, int64_t row_off, double dt, int32_t skip_flw) {
    Compiled cp;
    if (int st = compile_ops(c, blk.ops.data(), (int32_t)blk.ops.size(), cp)) return st;
    if (cp.unk != blk.unk || cp.has_edges != blk.has_edges || cp.has_diag != blk.has_diag)
        return set_error(VFVM_ERR_PATTERN_MISMATCH, "assemble: the scalar operators of unknown " + std::to_string(blk.unk) + " do not fit the built pattern");
    if (cp.has_time && !(dt > 0.0)) return set_error(VFVM_ERR_OPERATOR_USE_APPLY_TIME, "time operators need dt > 0 (OperatorUseApplyTime)");
    if (int st = prepare_scalar_fields(c, blk.ops.data(), (int32_t)blk.ops.size(), skip_flw)) return st;
    DevProgram p;
    build_program(c, cp, p);
    VFVM_HIP_TRY(blk.prog.alloc(1));
    VFVM_HIP_TRY(hipMemcpyAsync(blk.prog.p, &p, sizeof(p), hipMemcpyHostToDevice, c.stream));
    return launch_assemble_block(c, dt, c.mrow.p + row_off, blk.eslot.p, blk.dslot.p, blk.prog.p, c.b.p + row_off);
}

extern "C" int vfvm_assemble(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, double dt) {
    CTX_GUARD(ctx);
    if (int st = assemble_impl(c, ops, nops, dt)) return st;
    if (int st = sync_err(c)) return st;
    if (int st = ktime_flush(c)) return st;
    return clear_error();
}

extern "C" int vfvm_matrix_get(vfvm_ctx* ctx, double* vals, double* b) {
    CTX_GUARD(ctx);
    if (!c.has_matrix) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_matrix_get: nothing assembled");
    if (vals) VFVM_HIP_TRY(hipMemcpyAsync(vals, c.vals.p, c.nnz * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (b) VFVM_HIP_TRY(hipMemcpyAsync(b, c.b.p, c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (int st = sync_err(c)) return st;
    return clear_error();
}

// ------------------------------------------------------------------ solve
extern "C" void vfvm_solve_opts_default(vfvm_solve_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->tol = 1e-6;
    o->restart = 50;
    o->max_outer = 200;
    o->precision = VFVM_FP64;
    o->precond = VFVM_PRECOND_JACOBI;
    o->inner_tol = 1e-5;
    o->max_refine = 0;
}

extern "C" int vfvm_solve_device(vfvm_ctx* ctx, const vfvm_solve_opts* opts, double* x_dev, vfvm_solve_stats* stats) {
    CTX_GUARD(ctx);
    if (!opts || !x_dev) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_solve_device: NULL argument");
    if (!c.has_matrix) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_solve_device: nothing assembled");
    vfvm_solve_stats st;
    int r = gmres_solve(c, primal_view(c), *opts, x_dev, st);
    if (stats) *stats = st;
    if (r != VFVM_OK) return r;
    return clear_error();
}

extern "C" int vfvm_solve(vfvm_ctx* ctx, const vfvm_solve_opts* opts, double* x, vfvm_solve_stats* stats) {
    CTX_GUARD(ctx);
    if (!opts || !x) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_solve: NULL argument");
    if (!c.has_matrix) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_solve: nothing assembled");
    VFVM_HIP_TRY(c.x64.alloc(c.n));
    VFVM_HIP_TRY(hipMemcpyAsync(c.x64.p, x, c.n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    vfvm_solve_stats st;
    int r = gmres_solve(c, primal_view(c), *opts, c.x64.p, st);
    if (stats) *stats = st;
    VFVM_HIP_TRY(hipMemcpyAsync(x, c.x64.p, c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (int s2 = sync_err(c)) return s2;
    if (r != VFVM_OK) return r;
    return clear_error();
}

extern "C" int vfvm_residual(vfvm_ctx* ctx, const double* x, double* rel_residual) {
    CTX_GUARD(ctx);
    if (!x || !rel_residual) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_residual: NULL argument");
    if (!c.has_matrix) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_residual: nothing assembled");
    VFVM_HIP_TRY(c.x64.alloc(c.n));
    VFVM_HIP_TRY(hipMemcpyAsync(c.x64.p, x, c.n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    double bn = 0.0;
    if (int st = residual_norm(c, primal_view(c), c.x64.p, rel_residual, &bn)) return st;
    return clear_error();
}

extern "C" int vfvm_precond_apply(vfvm_ctx* ctx, const vfvm_solve_opts* opts, const double* r, double* z) {
    CTX_GUARD(ctx);
    if (!opts || !r || !z) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_precond_apply: NULL argument");
    if (!c.has_matrix) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_precond_apply: nothing assembled");
    if (int st = precond_apply(c, *opts, r, z)) return st;
    return clear_error();
}

extern "C" int vfvm_ilu_order(vfvm_ctx* ctx, int32_t* perm, int32_t* color_ptr, int32_t* ncolors) {
    CTX_GUARD(ctx);
    if (!c.has_pattern) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_ilu_order: no pattern built");
    if (int st = ilu_order(c, perm, color_ptr, ncolors)) return st;
    return clear_error();
}

extern "C" int vfvm_ilu_wait_limit(vfvm_ctx* ctx, int32_t polls, uint64_t* lapsed) {
    CTX_GUARD(ctx);
    if (int st = ilu_wait_limit(c, polls, lapsed)) return st;
    return clear_error();
}

extern "C" int vfvm_ilu_tiles(vfvm_ctx* ctx, int32_t* tile_ptr, int32_t* ntiles, int32_t* level) {
    CTX_GUARD(ctx);
    if (!c.has_pattern) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_ilu_tiles: no pattern built");
    if (int st = ilu_tiles(c, tile_ptr, ntiles, level)) return st;
    return clear_error();
}

// ------------------------------------------------------------------ drivers
int vfvm::initial_dirichlet(Ctx& c, const vfvm_op* ops, int32_t nops) {
    if (is_flow_list(ops, nops)) return flow_initial_dirichlet(c, ops, nops);
    return scalar_initial_dirichlet(c, ops, nops);
}

int vfvm::scalar_initial_dirichlet(Ctx& c, const vfvm_op* ops, int32_t nops) {
    // OpSclBndDirichlet::apply_initial for every Dirichlet operator (steady_base.rs:53-54)
    for (int32_t i = 0; i < nops; ++i) {
        if (ops[i].kind != VFVM_OP_SCL_BND_DIRICHLET) continue;
        const int32_t b = ops[i].geom_id;
        const int64_t off = c.g.bnd_ptr[b], cnt = c.g.bnd_ptr[b + 1] - off;
        if (cnt == 0) continue;
        DevBuf<int32_t> nid;
        VFVM_HIP_TRY(nid.upload(c.g.bnd_nid.data() + off, cnt, c.stream));
        hipLaunchKernelGGL(k_scatter_bnd, dim3(grid_for(cnt)), dim3(BLK), 0, c.stream, cnt, c.g.NP, nid.p, c.sbnd[ops[i].a].value.p, c.scl[ops[i].unk_id].value.p);
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    }
    return VFVM_OK;
}

namespace {

struct PicardSums {
    double d2 = 0.0, mo = 0.0, mn = 0.0, nf = 0.0, o2 = 0.0;
};
int picard_sums(Ctx& c, const double* xo, const double* xn, int64_t n, PicardSums& s) {
    const unsigned nb = grid_for(n);
    VFVM_HIP_TRY(c.red.grow(PIC_PARTS * (size_t)nb));
    hipLaunchKernelGGL(k_picard, dim3(nb), dim3(BLK), 0, c.stream, n, xo, xn, c.red.p);
    std::vector<double> h(PIC_PARTS * (size_t)nb);
    VFVM_HIP_TRY(hipMemcpyAsync(h.data(), c.red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    for (unsigned k = 0; k < nb; ++k) {
        const double* p = h.data() + PIC_PARTS * (size_t)k;
        s.d2 += p[0]; s.mo = std::max(s.mo, p[1]); s.mn = std::max(s.mn, p[2]); s.nf += p[3]; s.o2 += p[4];
    }
    return VFVM_OK;
}

// scaled Picard residual for one scalar field of n entries (scaling.rs:88-135): rms of the
// change over the field scale max(|x| over both states, the field's user scale, 1e-30); with
// scale_off the unscaled |x_new - x_old| / (|x_old| + 1e-10)
int picard_residual(Ctx& c, const double* xo, const double* xn, double* res, int64_t n, int32_t fid) {
    PicardSums s;
    if (int st = picard_sums(c, xo, xn, n, s)) return st;
    if (s.nf > 0.0) { *res = NAN; return VFVM_OK; }
    if (c.scale_off) { *res = std::sqrt(s.d2) / (std::sqrt(s.o2) + 1e-10); return VFVM_OK; }
    const double rms = std::sqrt(s.d2 / (double)std::max<int64_t>(n, 1));
    const double scale = std::max(std::max(std::max(s.mo, s.mn), c.user_scale_scl(fid)), 1e-30);
    *res = rms / scale;
    return VFVM_OK;
}

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

int picard_loop(Ctx& c, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts, double dt, int32_t max_iter,
                double tol, double damp, vfvm_driver_stats& ds, int ts, bool transient) {
    const double t_now = transient ? dt * (double)(ts + 1) : 0.0;  // end of the step (transient_base.rs:71)
    const int64_t n = c.n;
    SclDom& u = c.scl[c.unk];
    DevBuf<double>& xi = c.pic_xi;
    DevBuf<double>& xu = c.pic_xu;
    VFVM_HIP_TRY(xi.grow(n));
    VFVM_HIP_TRY(xu.grow(n));
    auto t0 = clk::now();
    // x_iter = current unknown values (_unkn_to_solu, steady_base.rs:137)
    VFVM_HIP_TRY(hipMemcpyAsync(xi.p, u.value.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    if (int st = call_update_hook(c, t_now)) return st;
    if (int st = assemble_impl(c, ops, nops, dt)) return st;
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    auto t1 = clk::now();
    ds.time_assembly += secs(t0, t1);
    for (int32_t iter = 0; iter < max_iter; ++iter) {
        auto ti0 = clk::now();
        VFVM_HIP_TRY(hipMemcpyAsync(xu.p, xi.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        vfvm_solve_stats st;
        if (int r = gmres_solve(c, primal_view(c), *opts, xu.p, st)) return r;
        ds.linear_iterations += st.iterations;
        auto ti1 = clk::now();
        ds.time_solve += secs(ti0, ti1);
        // damping into the unknown; re-assemble at the new iterate
        hipLaunchKernelGGL(k_damp, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, damp, xi.p, xu.p, u.value.p);
        if (int st1 = call_update_hook(c, t_now)) return st1;
        if (int st2 = assemble_impl(c, ops, nops, dt)) return st2;
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        ds.time_assembly += secs(ti1, clk::now());
        double res = 0.0;
        if (int st3 = picard_residual(c, xi.p, u.value.p, &res, n, c.unk)) return st3;
        ds.residual = res;
        ds.picard_iterations += 1;
        if (!std::isfinite(res))
            return set_error(transient ? VFVM_ERR_TRANSIENT_NON_FINITE_RESIDUAL : VFVM_ERR_STEADY_NON_FINITE_RESIDUAL,
                             "non-finite Picard residual at iteration " + std::to_string(iter) + (transient ? " of step " + std::to_string(ts) : ""));
        VFVM_HIP_TRY(hipMemcpyAsync(xi.p, u.value.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        if (res < tol) return VFVM_OK;
    }
    char buf[160];
    snprintf(buf, sizeof(buf), "Picard iteration did not converge in %d iterations (residual %.3e)", max_iter, ds.residual);
    return set_error(transient ? VFVM_ERR_TRANSIENT_DID_NOT_CONVERGE : VFVM_ERR_STEADY_DID_NOT_CONVERGE, buf);
}

}  // namespace

int vfvm::scalar_picard_residual(Ctx& c, const double* xo, const double* xn, double* res) {
    return picard_residual(c, xo, xn, res, c.n, c.unk);
}

int vfvm::scalar_field_residual(Ctx& c, const double* xo, const double* xn, int64_t n, int32_t fid, double* res) {
    return picard_residual(c, xo, xn, res, n, fid);
}

int vfvm::unscaled_residual(Ctx& c, const double* xo, const double* xn, int64_t n, double* res) {
    PicardSums s;
    if (int st = picard_sums(c, xo, xn, n, s)) return st;
    *res = s.nf > 0.0 ? NAN : std::sqrt(s.d2) / (std::sqrt(s.o2) + 1e-10);
    return VFVM_OK;
}

// ------------------------------------------------------------------ convergence scaling
// Variables::set_scldom_scale / set_vecdom_scale (vars.rs:203-213): a reference scale of an
// unknown field, a lower bound of its scale in the Picard residual and the adaptive error norm
extern "C" int vfvm_set_scldom_scale(vfvm_ctx* ctx, int32_t id, double scale) {
    CTX_GUARD(ctx);
    if (std::isnan(scale) || scale <= 0.0)
        return set_error(VFVM_ERR_VARIABLES_NON_POSITIVE_SCALE, "reference scale must be greater than zero (VariablesNonPositiveScale)");
    if (!check_scl(c, id)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_set_scldom_scale: id out of range");
    if ((int32_t)c.scale_scl.size() <= id) c.scale_scl.resize(id + 1, 0.0);
    c.scale_scl[id] = scale;
    return clear_error();
}

extern "C" int vfvm_set_vecdom_scale(vfvm_ctx* ctx, int32_t id, double scale) {
    CTX_GUARD(ctx);
    if (std::isnan(scale) || scale <= 0.0)
        return set_error(VFVM_ERR_VARIABLES_NON_POSITIVE_SCALE, "reference scale must be greater than zero (VariablesNonPositiveScale)");
    if (id < 0 || id >= (int32_t)c.vec.size()) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_set_vecdom_scale: id out of range");
    if ((int32_t)c.scale_vec.size() <= id) c.scale_vec.resize(id + 1, 0.0);
    c.scale_vec[id] = scale;
    return clear_error();
}

// Variables::scale_off (vars.rs:31)
extern "C" int vfvm_set_scale_off(vfvm_ctx* ctx, int32_t off) {
    CTX_GUARD(ctx);
    c.scale_off = off != 0;
    return clear_error();
}

int vfvm::eval_affine_closures(Ctx& c) {
    const int64_t N = c.g.N;
    for (const Ctx::VecAffine& a : c.vaff) {
        VecDom& f = c.vec[a.vec];
        hipLaunchKernelGGL(k_affine, dim3(grid_for(3 * N)), dim3(BLK), 0, c.stream, N, (const double*)a.f0.p, (const double*)a.df.p,
                           (const double*)c.scl[a.scl].value.p, f.value.p);
        VFVM_HIP_TRY(hipGetLastError());
        f.grad_valid = false;
    }
    return VFVM_OK;
}

int vfvm::call_update_hook(Ctx& c, double t) {
    if (int st = eval_affine_closures(c)) return st;
    if (!c.update_hook) return VFVM_OK;
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    c.update_hook(c.update_user, t);
    return VFVM_OK;
}

// Variables::add_vecdom_func with an affine closure of one scalar field (vars.rs:153):
// vec := f0 + df s, component-wise (3N SoA arrays), re-evaluated on the device wherever the
// drivers evaluate function fields (_update_func); df is what the flow adjoint differentiates
extern "C" int vfvm_vecdom_func_affine(vfvm_ctx* ctx, int32_t vec_id, int32_t scl_id, const double* f0, const double* df) {
    CTX_GUARD(ctx);
    if (vec_id < 0 || vec_id >= (int32_t)c.vec.size() || scl_id < 0 || scl_id >= (int32_t)c.scl.size())
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_vecdom_func_affine: field id out of range");
    if (c.vec[vec_id].kind != VFVM_FIELD_FUNCTION)
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_vecdom_func_affine: the vector field must be a function field");
    size_t k = 0;
    while (k < c.vaff.size() && c.vaff[k].vec != vec_id) ++k;
    if (!f0 || !df) {  // removes the closure
        if (k < c.vaff.size()) c.vaff.erase(c.vaff.begin() + k);
        return clear_error();
    }
    if (k == c.vaff.size()) c.vaff.emplace_back();
    Ctx::VecAffine& a = c.vaff[k];
    a.vec = vec_id;
    a.scl = scl_id;
    const int64_t N = c.g.N;
    VFVM_HIP_TRY(a.f0.upload(f0, 3 * N, c.stream));
    VFVM_HIP_TRY(a.df.upload(df, 3 * N, c.stream));
    if (int st = eval_affine_closures(c)) return st;
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_set_update_hook(vfvm_ctx* ctx, vfvm_update_hook hook, void* user) {
    CTX_GUARD(ctx);
    c.update_hook = hook;
    c.update_user = user;
    return clear_error();
}

extern "C" int vfvm_steady_solve(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts,
                                 int32_t max_iter, double tol, double damp, vfvm_driver_stats* stats) {
    CTX_GUARD(ctx);
    // steady_base.rs:21-23
    if (max_iter <= 0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_MAX_ITER, "max_iter must be positive");
    if (std::isnan(tol) || tol <= 0.0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TOLERANCE, "tol must be positive");
    if (std::isnan(damp) || damp <= 0.0 || damp > 1.0) return set_error(VFVM_ERR_SOLVER_INVALID_DAMPING, "damp must be in (0, 1]");
    if (!opts) return set_error(VFVM_ERR_INVALID_ARGUMENT, "opts is NULL");
    vfvm_driver_stats ds;
    std::memset(&ds, 0, sizeof(ds));
    if (int st = vfvm_pattern_build(ctx, ops, nops, nullptr)) return st;
    if (int st = call_update_hook(c, 0.0)) return st;  // _update_func before initial_dirichlet (steady_base.rs:50-54)
    if (int st = initial_dirichlet(c, ops, nops)) return st;
    int r = c.flow ? flow_picard(c, ops, nops, *opts, max_iter, tol, damp, ds)
                   : picard_loop(c, ops, nops, opts, 0.0, max_iter, tol, damp, ds, 0, false);
    if (stats) *stats = ds;
    if (r != VFVM_OK) return r;
    return clear_error();
}

extern "C" int vfvm_transient_solve(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts,
                                    double dt, int32_t num_ts, int32_t max_iter, double tol, double damp,
                                    vfvm_driver_stats* stats) {
    CTX_GUARD(ctx);
    // transient_base.rs:21-26
    if (std::isnan(dt) || dt <= 0.0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TIME_STEP, "dt must be positive");
    if (num_ts <= 0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_NUM_STEPS, "num_ts must be positive");
    if (max_iter <= 0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_MAX_ITER, "max_iter must be positive");
    if (std::isnan(tol) || tol <= 0.0) return set_error(VFVM_ERR_SOLVER_NON_POSITIVE_TOLERANCE, "tol must be positive");
    if (std::isnan(damp) || damp <= 0.0 || damp > 1.0) return set_error(VFVM_ERR_SOLVER_INVALID_DAMPING, "damp must be in (0, 1]");
    if (!opts) return set_error(VFVM_ERR_INVALID_ARGUMENT, "opts is NULL");
    vfvm_driver_stats ds;
    std::memset(&ds, 0, sizeof(ds));
    if (int st = vfvm_pattern_build(ctx, ops, nops, nullptr)) return st;
    // transient_base.rs:55-64: _update_func, the Dirichlet values at t = 0, _update_func again
    // at the initial state, and _update_prev of every field (the initial state is the previous
    // state of the first step)
    if (int st = call_update_hook(c, 0.0)) return st;
    if (int st = initial_dirichlet(c, ops, nops)) return st;
    if (int st = call_update_hook(c, 0.0)) return st;
    if (c.flow && !c.fs.transient)
        return set_error(VFVM_ERR_OPERATOR_UNSUPPORTED, "transient flow needs the velocity time operator (VEC_DOM_TIME) first");
    if (int st = update_prev_all(c)) return st;
    if (c.flow) {
        // TransientFlow (transient_flow.rs:248-303 with transient_base.rs:16-160): Picard per
        // step, then _update_prev
        int r = VFVM_OK;
        for (int32_t ts = 0; ts < num_ts && r == VFVM_OK; ++ts) {
            r = flow_picard(c, ops, nops, *opts, max_iter, tol, damp, ds, dt, ts);
            if (r == VFVM_OK) r = update_prev_all(c);
        }
        if (stats) *stats = ds;
        if (r != VFVM_OK) return r;
        return clear_error();
    }
    // scalar list: Picard per step, then _update_prev of every field (transient_base.rs:131,
    // :297-311), so a weight that the update hook changes enters the next step's
    // TIME_PREVIOUS terms with its value at the end of this step
    int r = VFVM_OK;
    for (int32_t ts = 0; ts < num_ts && r == VFVM_OK; ++ts) {
        r = picard_loop(c, ops, nops, opts, dt, max_iter, tol, damp, ds, ts, true);
        if (r == VFVM_OK) r = update_prev_all(c);
    }
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    if (stats) *stats = ds;
    if (r != VFVM_OK) return r;
    return clear_error();
}

// ------------------------------------------------------------------ timing
extern "C" int vfvm_timing_enable(vfvm_ctx* ctx, int32_t enable) {
    CTX_GUARD(ctx);
    c.timing_mask = (uint32_t)enable;
    return clear_error();
}
extern "C" int vfvm_timing_get(vfvm_ctx* ctx, int32_t kclass, int64_t* launches, double* total_ms, double* bytes) {
    CTX_GUARD(ctx);
    if (kclass < 0 || kclass >= VFVM_K_COUNT) return set_error(VFVM_ERR_INVALID_ARGUMENT, "kernel class out of range");
    if (int st = ktime_flush(c)) return st;
    if (launches) *launches = c.kt[kclass].launches;
    if (total_ms) *total_ms = c.kt[kclass].ms;
    if (bytes) *bytes = c.kt[kclass].bytes;
    return clear_error();
}
extern "C" int vfvm_timing_reset(vfvm_ctx* ctx) {
    CTX_GUARD(ctx);
    if (int st = ktime_flush(c)) return st;
    for (auto& k : c.kt) k = KTimer();
    return clear_error();
}
