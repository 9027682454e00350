// Device context: geometry, fields, pattern, matrix and solver workspace in HBM.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"
#include "mesh.hpp"

namespace vfvm {

#define VFVM_HIP_TRY(expr)                                                                    \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return set_error(VFVM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Owning device allocation.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; return *this; }
    ~DevBuf() { reset(); }
    void reset() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    hipError_t alloc(size_t count) {
        if (count == n && p) return hipSuccess;
        reset();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e == hipSuccess) n = count; else p = nullptr;
        return e;
    }
    // keeps the allocation when it already holds `count` entries (scratch reused across calls)
    hipError_t grow(size_t count) {
        if (p && n >= count) return hipSuccess;
        return alloc(count);
    }
    hipError_t upload(const T* h, size_t count, hipStream_t s) {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0) return e;
        return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

// Device copy of the dual geometry (SoA, CSR edge order).
struct DevGeom {
    int64_t N = 0, NP = 0, E = 0, S = 0;
    int32_t num_bnd = 0;
    DevBuf<int32_t> row_ptr, ngh, erow;       // erow[e] = row of directed edge e
    DevBuf<double> dist, area, norm, skew, gw; // norm/skew/gw: 3E SoA
    DevBuf<double> vol, cskew;                 // cskew: 3N SoA
    DevBuf<int32_t> bsf_ptr, bsf_vj, bsf_vk;
    DevBuf<double> bsf_area, bsf_wgt, bsf_norm; // bsf_wgt, bsf_norm: 3S SoA
    DevBuf<int32_t> vbnd, bidx;                // vertex -> boundary (-1: point), index in boundary
    // operator class of every vertex: the points of domain d are class d, boundary b is class
    // num_dom + b, interface k's side-a / side-b node copies are classes cls_itf0 + 2k / + 2k + 1
    DevBuf<int32_t> vcls;
    std::vector<int32_t> h_vcls;
    int32_t num_dom = 1, num_itf = 0, cls_itf0 = 1;
    bool multi = false;                        // joined mesh (several domains or interfaces)
    // interface nodes over all interfaces (q = 0 .. K-1): the two copies and the interface
    int64_t K = 0;
    DevBuf<int32_t> itf_va, itf_vb, itf_of;
    std::vector<int32_t> h_itf_va, h_itf_vb, h_itf_of;
    DevBuf<int32_t> rev;                       // rev[e] = directed edge (n, v) of e = (v, n); built on demand
    std::vector<int64_t> bnd_ptr;              // host copy
    std::vector<int32_t> bnd_nid;
    std::vector<int32_t> h_row_ptr, h_ngh;     // host copies for the pattern build
    std::vector<double> h_pos;                 // vertex coordinates (3N AoS): row order of the ILU(0) colours
    // pressure-gradient coefficients (flow operators), when the mesh has them
    int64_t G = 0;
    DevBuf<int32_t> gc_ptr, gc_vid;
    DevBuf<double> gc_c;                       // 3G SoA
    std::vector<int32_t> h_gc_ptr, h_gc_vid;
    // transpose of the coefficient lists, built on demand (flow adjoint): the slots q with
    // gc_vid[q] = u are gct_slot[gct_ptr[u] .. gct_ptr[u+1]), owner vertex gct_own (ascending)
    DevBuf<int32_t> gct_ptr, gct_slot, gct_own;
};

struct SclDom {
    int kind = 0;
    double cnst = 0.0;
    bool grad_valid = false;
    DevBuf<double> value, grad, value_prev;  // grad: 3N SoA
};
struct VecDom {
    int kind = 0;
    bool grad_valid = false;
    DevBuf<double> value, grad;  // value 3N SoA, grad 9N SoA (u_x,u_y,u_z,v_x,...)
    DevBuf<double> value_prev;   // 3N: values of the previous time step
};
struct SclBnd {
    int32_t bnd = 0;
    DevBuf<double> value;  // per node of the boundary
};
struct SclFlw {
    int32_t vel = -1, pres = -1, den = -1, visc = -1, fce = -1;
    DevBuf<double> flux;  // E
    DevBuf<double> rc;    // N: V / a_P (Rhie-Chow coefficient)
    DevBuf<double> mom;   // 3N SoA: rc (grad p - fce_c)
};
struct VecBnd {
    int32_t bnd = 0;
    DevBuf<double> value;  // 3 x (nodes of the boundary) SoA
};

// One operator as the assembly kernel reads it.
struct DevOp {
    int32_t kind;
    int32_t has_diff;
    const double* f_val;   // coefficient field value (or sclbnd values)
    const double* f_grad;  // coefficient field gradient (3N SoA)
    const double* u_grad;  // unknown gradient (production)
    const double* u_prev;  // unknown previous value (time)
    const double* flux;    // face mass flux (advection)
    const double* w_val;   // weight field value and gradient (advection: m)
    const double* w_grad;
    const double* r_val;   // density value and gradient (advection: rho)
    const double* r_grad;
    int32_t side;          // interface advection: flowing side (0: a, 1: b)
    int32_t pad;
};
// class d = points of domain d, then one class per boundary, then two per interface
constexpr int MAX_CLASSES = 32;
// columns of an interface row (per interface class): the faces of the side-a / side-b copy of
// the node (ngh(a_v) / ngh(b_v)) and the two copies themselves
constexpr uint8_t ITF_EA = 1, ITF_EB = 2, ITF_CA = 4, ITF_CB = 8;
constexpr int MAX_OPS = 8;
struct DevProgram {
    int32_t nops[MAX_CLASSES];
    DevOp ops[MAX_CLASSES][MAX_OPS];
};

// Multicolour ILU(0) ordering of one pattern, in colour-major storage: row t of the
// permuted system is row perm[t] of the matrix, rows sorted by (colour, Morton code of the
// vertex); entries [L | diag | U] with permuted columns; src[k] = CSR slot of permuted entry
// k.  The factor is stored split by triangle, [L (nnz_l) | U (nnz_u) | diag (n)] with the
// index arrays lptr / lcol and uptr / ucol, so that each sweep streams only its triangle;
// lpos / upos give the entries' positions in the colour-major CSR.
struct IluData {
    DevBuf<int32_t> perm, prp, src, pcol, pdiag;
    DevBuf<int32_t> lptr, lcol, lpos, uptr, ucol, upos;
    int64_t nnz_l = 0, nnz_u = 0;
    std::vector<int32_t> h_color_ptr;
    // tiled order (the rows of a tile contiguous, its levels inside): the first row of every level
    // of every tile, the factorisation's row groups, and the sweeps' roles (all tile sweeps of
    // one application in one launch): per role 16 ints {r0, r1, L entries begin, end, U entries
    // begin, end, tile, kind, dependencies begin, end, 0 ...}, the dependency lists, and the
    // launches' ticket counter, lapsed-wait count and per-tile completion flags
    bool tiled = false;
    DevBuf<int32_t> lev_row, fac_rows, trole, tdep;
    DevBuf<unsigned long long> tsync;
    std::vector<int32_t> h_fac_ptr, h_tile_row, h_lev_row;
    std::vector<int32_t> h_prp;  // host copy of the permuted CSR's row pointers
    int32_t nrole = 0;
    double sweep_rows = 0.0, sweep_lent = 0.0, sweep_uent = 0.0, fused_rows = 0.0;  // algorithmic byte counts
    int32_t nlev = 0, rmax = 0, emax_l = 0, emax_u = 0;
    bool ready = false;
    uint64_t generation = 0;  // bumped by every set-up (graphs bake the layout in)
    // polls of a completion flag before a tile sweep's wait gives up (< 0: every wait gives up
    // at once, for tests of that path)
    int32_t wait_limit = 1 << 17;
};

// One linear system as the solver reads it (non-owning).
struct SysView {
    int64_t n = 0, nnz = 0;
    const int32_t* row = nullptr;
    const int32_t* col = nullptr;
    const int32_t* dslot = nullptr;
    const double* vals = nullptr;
    const double* b = nullptr;
    const std::vector<int32_t>* h_row = nullptr;
    const std::vector<int32_t>* h_col = nullptr;
    IluData* ilu = nullptr;
    // coupled flow system: 0 generic CSR, 1 structured rows in the comp-major order,
    // 2 structured rows with the vectors vertex-interleaved (entry (c, v) at (4 + flow_ns) v + c)
    int flow_layout = 0;
    int flow_ns = 0;  // scalar unknowns coupled to the flow (rows (4 + k) N + v), at most 2
    // generic CSR whose rows, columns and vectors are numbered vertex-interleaved (unknown c of
    // vertex v at (n / N) v + c): block Jacobi reads and writes its blocks there
    bool interleaved = false;
    // most entries held by one chunk of SPMV_ROWS consecutive rows (0: the per-row SpMV kernel)
    int32_t spmv_emax = 0;
};

// Checkpoint tape of the adjoint forward pass: the operator list, step data and the
// states T^0, T^K, T^2K, ... (K = checkpoint interval) kept in HBM.
struct AdjTape {
    bool valid = false;
    std::vector<vfvm_op> ops;
    double dt = 0.0;
    int32_t num_ts = 0;  // 0: steady
    int32_t every = 0;
    vfvm_solve_opts opts{};
    DevBuf<double> ckpt;   // (num_ts / every + 1) states of n entries
    DevBuf<double> seg;    // (every + 1) states T^first .. T^(first + count) of one segment
    int32_t seg_first = -1, seg_count = 0;
    DevBuf<double> final_state;
    vfvm_adjoint_stats stats{};
    // coupled flow system (vfvm_flow_adjoint_forward): the Picard controls of every step,
    // replayed exactly by the backward pass, and the forward pass's own totals
    bool flow = false;
    int32_t picard_max = 0;
    double picard_tol = 0.0, picard_damp = 1.0;
    int32_t picard_iterations = 0;
    double objective = 0.0;
};

// Captured launches of one restart cycle (one executable graph per chunk of steps).
struct CycleGraphs {
    std::array<uint64_t, 22> key{};
    std::vector<hipGraphExec_t> chunks;
    uint64_t last_use = 0;
};

// Coupled velocity-pressure system of a flow operator list (flow.hip).
struct FlowSys {
    int32_t vel = -1, pres = -1, den = -1, visc = -1, fce = -1, flw = -1;
    std::vector<int32_t> vbnd_val;  // per boundary: vecbnd id of its velocity condition
    int64_t pref_row = -1;          // row of the reference pressure (-1: none)
    double pref_val = 0.0;
    int32_t gmax = 0, rmax = 0;     // largest grad_coeff list and 3-ring of a continuity row
    int32_t lpr = 16;               // lanes per momentum row
    int64_t spmv_idx = 0;           // column indices the structured SpMV reads per product
    bool transient = false;         // TransientFlow list (velocity time operator first)
    double dt = 0.0;
    std::vector<int32_t> blk_unk;   // scalar unknowns coupled in (rows (4 + k) N + v)
    std::vector<std::vector<vfvm_op>> blk_ops;  // their operators, in list order
};

// Block-Jacobi inverse blocks (b x b per vertex, entry (i, j) of vertex v at (i b + j) N + v).
struct BlockJacobi {
    int32_t nb = 0;
    DevBuf<double> inv64;
    DevBuf<float> inv32;
};

// A scalar unknown's rows inside a coupled flow system (heat, species): its operators in list
// order, its slot maps into the coupled CSR, its device operator program.
struct ScalarBlock {
    int32_t unk = -1;
    std::vector<vfvm_op> ops;
    std::vector<uint8_t> has_edges, has_diag;  // per vertex class, as the pattern was built
    DevBuf<int32_t> eslot, dslot;
    DevBuf<DevProgram> prog;
};

// A system assembled by the caller from triplets (LinearBase): the triplets of CSR slot s are
// trip[sptr[s] .. sptr[s+1]) in list order; tv holds the uploaded triplet values.
struct LinearSys {
    bool active = false;
    int64_t ntrip = 0;
    DevBuf<int64_t> sptr, trip;
    DevBuf<double> tv;
};

struct KTimer {
    int64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DevGeom g;
    bool has_mesh = false;
    std::vector<SclDom> scl;
    std::vector<VecDom> vec;
    std::vector<SclBnd> sbnd;
    std::vector<SclFlw> flw;
    std::vector<VecBnd> vbndv;

    // pattern
    bool has_pattern = false;
    int32_t unk = -1;
    int64_t n = 0, nnz = 0;
    std::vector<int32_t> h_mrow, h_mcol;
    DevBuf<int32_t> mrow, mcol, eslot, dslot;
    // interface rows (joined meshes): per interface node q the slots (a_v, a_v), (a_v, b_v) of
    // the side-a row and (b_v, a_v), (b_v, b_v) of the side-b row (-1: absent), 4 per node;
    // eslot of a face of either copy points into the side-b row
    DevBuf<int32_t> islot;
    std::vector<vfvm_op> pattern_ops;
    DevBuf<DevProgram> prog;
    DevBuf<double> vals, b;
    bool has_matrix = false;

    LinearSys lin;  // set when the pattern came from vfvm_linear_update_sparse
    IluData ilu;  // ILU(0) ordering of the pattern above
    BlockJacobi bjac;

    // coupled flow system (when the pattern was built from flow operators)
    bool flow = false;
    FlowSys fs;
    std::vector<ScalarBlock> sblk;

    // convergence scaling (Variables::scale_scldom / scale_vecdom / scale_off, vars.rs:28-31):
    // user reference scales per field id (0: none), lower bounds of the field scales of the
    // Picard residual and the adaptive error norm (scaling.rs:88-111); scale_off selects the
    // unscaled residual |x_new - x_old| / |x_old| (scaling.rs:120-121)
    std::vector<double> scale_scl, scale_vec;
    bool scale_off = false;
    double user_scale_scl(int32_t id) const { return id >= 0 && id < (int32_t)scale_scl.size() ? scale_scl[id] : 0.0; }
    double user_scale_vec(int32_t id) const { return id >= 0 && id < (int32_t)scale_vec.size() ? scale_vec[id] : 0.0; }

    // host closure of function fields that depend on unknowns (Variables::_update_func),
    // called before every assembly of the drivers
    vfvm_update_hook update_hook = nullptr;
    void* update_user = nullptr;
    // device closures of vector function fields affine in one scalar field,
    // vec = f0 + df s (Variables::add_vecdom_func for an affine func, vars.rs:153), evaluated
    // at the same points as the host hook, before it
    struct VecAffine {
        int32_t vec = -1, scl = -1;
        DevBuf<double> f0, df;  // 3N SoA each
    };
    std::vector<VecAffine> vaff;

    // transposed system A^T (adjoint solves): CSR of the columns of A, tsrc[k] = slot of
    // entry k in vals; values gathered on demand
    bool has_transpose = false;
    bool t_interleaved = false;  // A^T numbered vertex-interleaved (coupled flow systems)
    std::vector<int32_t> h_trow, h_tcol;
    DevBuf<int32_t> trow, tcol, tsrc, tdslot;
    DevBuf<double> tvals, tb;
    IluData tilu;

    // solver workspace (shared by both systems)
    DevBuf<double> V64, w64, z64, x64, r64, dinv64, avals64, part, scal;
    DevBuf<float> V32, w32, avals32;
    DevBuf<double> ilu64, ap64, bp64, xp64;  // ILU factor (split); A, b, x in the colour-major order
    DevBuf<float> ilu32;
    DevBuf<double> iluf64;  // factor in the colour-major [L | diag | U] layout while it is computed
    DevBuf<float> iluf32;
    double* h_pinned = nullptr;  // small pinned readback buffer
    // driver scratch, kept across Picard iterations and time steps: per-block partial sums of
    // the residual and statistics reductions, and the Picard iterates
    DevBuf<double> red;
    DevBuf<int> red_i;
    DevBuf<double> pic_xi, pic_xu, pic_xd;

    AdjTape tape;

    // cycle graphs (krylov.hip)
    bool graphs_enabled = true;
    uint64_t graph_clock = 0;
    std::vector<CycleGraphs> graphs;
    // completion of the control-word readback of the two most recent cycle-graph chunks
    hipEvent_t ctrl_ev[2] = {nullptr, nullptr};

    // timing
    uint32_t timing_mask = 0;  // bit k: time kernel class k
    std::array<KTimer, VFVM_K_COUNT> kt;
    std::vector<hipEvent_t> ev_pool;
    struct Pending { int kclass; double bytes; hipEvent_t a, b; int step; };
    int cur_step = -1;
    std::vector<Pending> pending;
    size_t ev_next = 0;
};

// timing helpers (no-ops unless timing is on)
void ktime_begin(Ctx& c, int kclass, hipEvent_t* a);
void ktime_end(Ctx& c, int kclass, double bytes, hipEvent_t a);
int ktime_flush(Ctx& c);

// kernels (kernels_asm.hip)
int launch_grad(Ctx& c, const double* value, double* grad);
int launch_grad_vec(Ctx& c, const double* value3, double* grad9);
int launch_flux_prescribed(Ctx& c, SclFlw& f);
int launch_assemble(Ctx& c, double dt);
int launch_assemble_block(Ctx& c, double dt, const int32_t* mrow, const int32_t* eslot, const int32_t* dslot,
                          const DevProgram* prog, double* b);

// solver (krylov.hip)
SysView primal_view(Ctx& c);
SysView transpose_view(Ctx& c);
int gmres_solve(Ctx& c, const SysView& s, const vfvm_solve_opts& o, double* x_dev, vfvm_solve_stats& st);
// An inner solve of an outer iteration: x from 0, the FP32 matrix and the preconditioner of the
// previous solve reused when reuse_setup (the caller guarantees the same values), and no residual
// after the last allowed cycle (status VFVM_ERR_LINEAR_DID_NOT_CONVERGE, rel_residual -1)
int gmres_solve_inner(Ctx& c, const SysView& s, const vfvm_solve_opts& o, double* x_dev, vfvm_solve_stats& st, bool reuse_setup);
// The same solve (FP64, Jacobi or no preconditioner) with its whole workspace carved from the
// caller's device region of `entries` doubles: no allocation, eager launches, host-clock timing.
// gmres_noalloc_entries gives the doubles a system needs.
int gmres_solve_noalloc(Ctx& c, const SysView& s, const vfvm_solve_opts& o, double* x_dev, vfvm_solve_stats& st,
                        double* scratch, int64_t entries);
int64_t gmres_noalloc_entries(int64_t n, int64_t nnz, int m, bool jacobi, bool interleaved);
// Looks up the device address of the current device's no-allocation scratch array (api_noalloc.hip)
int noalloc_setup(int device);
// host-side preconditioner ordering of a pattern (ILU(0) colouring), done once per pattern
int solver_setup(Ctx& c, const SysView& s, const vfvm_solve_opts& o);
// z = M^-1 r for the assembled matrix (host arrays); the ILU(0) row order
int precond_apply(Ctx& c, const vfvm_solve_opts& o, const double* r_host, double* z_host);
int ilu_order(Ctx& c, int32_t* perm, int32_t* color_ptr, int32_t* ncolors);
int ilu_tiles(Ctx& c, int32_t* tile_ptr, int32_t* ntiles, int32_t* level);
int ilu_wait_limit(Ctx& c, int32_t polls, uint64_t* lapsed);
int residual_norm(Ctx& c, const SysView& s, const double* x_dev, double* rel, double* bnorm);

// flow operators (flow.hip)
bool is_flow_list(const vfvm_op* ops, int32_t nops);
int flow_pattern_build(Ctx& c, const vfvm_op* ops, int32_t nops);
int flow_assemble(Ctx& c, const vfvm_op* ops, int32_t nops);
int flow_update(Ctx& c, int32_t flw);
int flow_initial_dirichlet(Ctx& c, const vfvm_op* ops, int32_t nops);
int flow_picard(Ctx& c, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts& o, int32_t max_iter, double tol,
                double damp, vfvm_driver_stats& ds, double dt = 0.0, int ts = 0);
int flow_assemble_dt(Ctx& c, const vfvm_op* ops, int32_t nops, double dt);
int update_prev_all(Ctx& c);
// x = [vel | pres] and back (the unknowns of the flow system)
int flow_pack(Ctx& c, double* x);
int flow_unpack(Ctx& c, const double* x);
// scaled Picard residual of the flow fields between two states (scaling.rs:113-135), and
// the field statistics of two states: [0] sum |du|^2, [1] max |vel|, [2] max |p|, [3] min p,
// [4] max p, [5] non-finite entries of xn, [6] max |rho|
int flow_residual(Ctx& c, const double* xo, const double* xn, double* res);
int flow_stats(Ctx& c, const double* xo, const double* xn, double out[7]);
// scaled Picard residual of the scalar unknown (api_ctx.hip)
int scalar_picard_residual(Ctx& c, const double* xo, const double* xn, double* res);
// scaled residual of one scalar field of n entries (scale lower bound: field fid's user scale)
int scalar_field_residual(Ctx& c, const double* xo, const double* xn, int64_t n, int32_t fid, double* res);
// |x_new - x_old|_2 / (|x_old|_2 + 1e-10) over n entries, NaN if x_new is not finite (the
// residual with scale_off, scaling.rs:116-122)
int unscaled_residual(Ctx& c, const double* xo, const double* xn, int64_t n, double* res);
// scalar operator lists inside a coupled system (api_ctx.hip)
int prepare_scalar_fields(Ctx& c, const vfvm_op* ops, int32_t nops, int32_t skip_flw);
int scalar_classes(Ctx& c, const vfvm_op* ops, int32_t nops, int32_t* unk, std::vector<uint8_t>& has_edges,
                   std::vector<uint8_t>& has_diag, bool* has_time);
int scalar_block_assemble(Ctx& c, ScalarBlock& blk, int64_t row_off, double dt, int32_t skip_flw);
int scalar_initial_dirichlet(Ctx& c, const vfvm_op* ops, int32_t nops);
// evaluates the device closures of affine function fields, then calls the host update hook
// (if any) with time t
int call_update_hook(Ctx& c, double t);
int eval_affine_closures(Ctx& c);
// y = A x over the view's CSR in FP64 (krylov.hip)
int spmv_apply(Ctx& c, const SysView& s, const double* x, double* y);
// transposed pattern of the context's system and the reverse edge map (adjoint.hip); with
// interleaved, A^T is numbered vertex-interleaved (unknown c of vertex v at S v + c, S = n / N)
int build_transpose(Ctx& c, bool interleaved = false);
int build_rev(Ctx& c);

// drivers and adjoint set-up shared across translation units (api_ctx.hip)
int assemble_impl(Ctx& c, const vfvm_op* ops, int32_t nops, double dt);
int initial_dirichlet(Ctx& c, const vfvm_op* ops, int32_t nops);
int sync_err(Ctx& c);

}  // namespace vfvm

struct vfvm_ctx {
    vfvm::Ctx c;
};

#define CTX_GUARD(ctx)                                                           \
    if (!(ctx)) return set_error(VFVM_ERR_INVALID_ARGUMENT, "context is NULL"); \
    vfvm::Ctx& c = (ctx)->c;                                                     \
    VFVM_HIP_TRY(hipSetDevice(c.device))
