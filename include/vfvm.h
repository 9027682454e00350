/*
 * vfvm.h — C ABI of the MI355X-native vfvm3d hot path.
 *
 * Scope: the per-Delaunay-edge Voronoi-face assembly of the scalar transport
 * operators into a fixed-pattern CSR matrix, and the restarted GMRES that solves
 * it, with the steady/transient Picard drivers that call them.  Every entry point
 * returns an int status (VFVM_OK = 0); the message of the last failure on the
 * calling thread is returned by vfvm_last_error().
 *
 * Plain pointers and sizes only.  Host arrays passed in are copied; the caller
 * keeps ownership.  A context owns one HIP stream on one device and is not
 * thread-safe: use one context per host thread.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   Tetrahedron::update_tet            src/base/geom_tet.rs:58      -> vfvm_mesh_from_tets / vfvm_mesh_lattice
 *                                                                      (dual on a GPU: vfvm_mesh_from_tets_dev / vfvm_mesh_lattice_dev)
 *   Tetrahedron::grad_coeff            src/base/geom_tet.rs:295     -> vfvm_mesh_build_grad_coeff (_dev on a GPU)
 *   write_{scl,vec}{dom,bnd}_csv       src/base/write_csv.rs:5-100  -> vfvm_write_csv
 *   write_{scl,vec}{dom,bnd}_vtu       src/base/write_vtu.rs        -> vfvm_write_vtu
 *   Variables::add_scldom_* / add_sclbnd_* / add_sclflw / add_vecdom_*
 *                                      src/base/vars.rs:80-183      -> vfvm_scldom_add / vfvm_sclbnd_add / vfvm_vecdom_add / vfvm_sclflw_add
 *   OperatorBase::apply / apply_time   src/operator/operator_base.rs:5-7,
 *                                      src/operator/core/opscl_dom_time.rs:35 -> vfvm_op descriptors
 *   LinearBase::update_sparse          src/linear/linear_base.rs:7  -> vfvm_pattern_build
 *   SteadyBase::assemble_matrix + LinearBase::build_matrix
 *                                      src/steady/steady_base.rs:61-66, src/linear/linear_gmres.rs:66-74 -> vfvm_assemble
 *   LinearBase::solve_matrix           src/linear/linear_gmres.rs:76-84 -> vfvm_solve
 *   SteadyBase::solve                  src/steady/steady_base.rs:16-135 -> vfvm_steady_solve
 *   TransientBase::solve               src/transient/transient_base.rs:16-160 -> vfvm_transient_solve
 *   SteadyFlow (flow operator list)    src/steady/core/steady_flow.rs:133-303 -> vfvm_op kinds 40-62 with
 *                                      vfvm_pattern_build / vfvm_assemble / vfvm_steady_solve
 *   flow_helper::update_flow           src/operator/core/flow_helper.rs:3-179 -> vfvm_update_flow
 *   Variables::add_vecbnd_{cnst,func}  src/base/vars.rs:169-184           -> vfvm_vecbnd_add
 *   TransientAdaptive::solve_adaptive_with src/transient/transient_adaptive.rs:81-350 -> vfvm_transient_adaptive
 *   LinearBase::{update_sparse, build_matrix, solve_matrix} for caller-assembled triplets
 *                                      src/linear/linear_base.rs:5-10, linear_gmres.rs:53-84
 *                                      -> vfvm_linear_update_sparse / vfvm_linear_build_matrix / vfvm_linear_solve_matrix
 *   Variables::set_scldom_scale / set_vecdom_scale / scale_off
 *                                      src/base/vars.rs:28-31,203-213 -> vfvm_set_scldom_scale / vfvm_set_vecdom_scale / vfvm_set_scale_off
 */
#ifndef VFVM_H
#define VFVM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirror the VFVChemError variants on this path, src/base/error.rs:5) ---- */
enum vfvm_status {
    VFVM_OK = 0,
    VFVM_ERR_INVALID_ARGUMENT = 1,          /* null pointer, size or id out of range            */
    VFVM_ERR_HIP = 2,                       /* HIP runtime error (message holds hipGetErrorString) */
    VFVM_ERR_NO_EXTENSION = 3,              /* device code not loadable                          */
    VFVM_ERR_TETRA_DEGENERATE_INPUT = 10,   /* TetraDegenerateInput   geom_tet.rs:88              */
    VFVM_ERR_TETRA_POINTS_ON_BOUNDARY = 11, /* TetraPointsOnBoundary  geom_tet.rs:120             */
    VFVM_ERR_TETRA_SINGULAR_GRADIENT = 12,  /* TetraSingularGradient  geom_tet.rs:284             */
    VFVM_ERR_OPERATOR_NON_UNKNOWN = 20,     /* OperatorNonUnknownInLinearSystem operator_base.rs:15 */
    VFVM_ERR_OPERATOR_USE_APPLY_TIME = 21,  /* OperatorUseApplyTime   opscl_dom_time.rs:66         */
    VFVM_ERR_OPERATOR_UNSUPPORTED = 22,     /* operator combination outside the built path        */
    VFVM_ERR_PATTERN_MISMATCH = 23,         /* assembly ops do not match the built pattern        */
    VFVM_ERR_LINEAR_NON_POSITIVE_TOLERANCE = 30, /* LinearGmresNonPositiveTolerance linear_gmres.rs:29 */
    VFVM_ERR_LINEAR_NON_POSITIVE_RESTART = 31,   /* LinearGmresNonPositiveRestart   linear_gmres.rs:32 */
    VFVM_ERR_LINEAR_NON_POSITIVE_MAX_OUTER = 32, /* LinearGmresNonPositiveMaxOuter  linear_gmres.rs:35 */
    VFVM_ERR_LINEAR_DID_NOT_CONVERGE = 33,       /* LinearGmresDidNotConverge       linear_gmres.rs:81 */
    VFVM_ERR_LINEAR_NON_FINITE = 34,             /* breakdown to NaN/Inf inside the Krylov solve      */
    VFVM_ERR_SOLVER_NON_POSITIVE_MAX_ITER = 40,  /* SolverNonPositiveMaxIter   steady_base.rs:21      */
    VFVM_ERR_SOLVER_NON_POSITIVE_TOLERANCE = 41, /* SolverNonPositiveTolerance steady_base.rs:22      */
    VFVM_ERR_SOLVER_INVALID_DAMPING = 42,        /* SolverInvalidDamping       steady_base.rs:23      */
    VFVM_ERR_SOLVER_NON_POSITIVE_TIME_STEP = 43, /* SolverNonPositiveTimeStep  transient_base.rs:21   */
    VFVM_ERR_SOLVER_NON_POSITIVE_NUM_STEPS = 44, /* SolverNonPositiveNumTimeSteps transient_base.rs:22 */
    VFVM_ERR_STEADY_DID_NOT_CONVERGE = 45,       /* SteadyDidNotConverge       steady_base.rs:116     */
    VFVM_ERR_STEADY_NON_FINITE_RESIDUAL = 46,    /* SteadyNonFiniteResidual    steady_base.rs:101     */
    VFVM_ERR_TRANSIENT_DID_NOT_CONVERGE = 47,    /* TransientDidNotConverge    transient_base.rs:127  */
    VFVM_ERR_TRANSIENT_NON_FINITE_RESIDUAL = 48, /* TransientNonFiniteResidual transient_base.rs:119  */
    VFVM_ERR_WRITE_CSV = 50,                     /* WriteCsv                   error.rs:116            */
    VFVM_ERR_WRITE_VTU = 51,                     /* WriteVtu                   error.rs:122            */
    VFVM_ERR_VARIABLES_NON_POSITIVE_SCALE = 60   /* VariablesNonPositiveScale  vars.rs:204,210         */
};

/* Message of the last failing call on this thread ("" if none). */
const char* vfvm_last_error(void);
/* Library version string. */
const char* vfvm_version(void);

/* =============================== mesh (host) =============================== */

typedef struct vfvm_mesh vfvm_mesh;

typedef struct vfvm_mesh_info {
    int64_t num_pts;      /* interior points (vid < num_pts)                         */
    int64_t num_nds;      /* boundary nodes  (vid = num_pts + nid)                   */
    int64_t num_tets;
    int64_t num_edges;    /* directed Delaunay edges (CSR length)                    */
    int64_t num_subfaces; /* boundary sub-faces                                      */
    int32_t num_bnd;      /* boundary groups                                         */
    int32_t max_degree;
    int64_t num_grad_coeff; /* pressure-gradient coefficients (0 until built)         */
    int32_t num_dom;      /* domains (1 unless joined by vfvm_mesh_join)                 */
    int32_t num_itf;      /* interfaces                                                  */
    int64_t num_itf_nodes; /* interface nodes over all interfaces                         */
} vfvm_mesh_info;

/* Jittered unit-cube lattice of n^3 vertices (h = 1/(n-1)), jitter uniform in
 * [-jitter*h, +jitter*h] per free axis from a SplitMix64 hash of (seed,i,j,k,axis),
 * quantised to multiples of 2^-22; exact Delaunay tetrahedralisation; Voronoi dual.
 * Boundaries (in id order): "x0","x1","y0","y1","z0","z1"; edge and corner nodes
 * belong to the first face of that order that contains them.
 * Replaces Variables::new_bounds (src/base/vars.rs:45) + Tetrahedron::update_tet
 * (src/base/geom_tet.rs:58) for the lattice meshes. */
int vfvm_mesh_lattice(int32_t n, uint64_t seed, double jitter, vfvm_mesh** out);

/* Mesh from caller-supplied vertices and tetrahedra (any orientation); builds the
 * dual.  pos: 3*(num_pts+num_nds) AoS.  tets: 4*num_tets vids.  Boundaries:
 * bnd_ptr[num_bnd+1] offsets into bnd_nid (node ids, each node on one boundary).
 * Replaces Tetrahedron::update_tet (src/base/geom_tet.rs:58-350). */
int vfvm_mesh_from_tets(int64_t num_pts, int64_t num_nds, const double* pos,
                        int64_t num_tets, const int32_t* tets,
                        int32_t num_bnd, const int64_t* bnd_ptr, const int32_t* bnd_nid,
                        vfvm_mesh** out);

/* The same two constructors with the Voronoi dual built on GPU `device` (topology,
 * face pieces, per-edge and per-vertex sums, least-squares weights; the boundary sub-faces
 * on the host): every array is bitwise equal to the host build's.  Replaces
 * Tetrahedron::update_tet (src/base/geom_tet.rs:58-293) like the host build; Delaunay
 * stays on the host. */
int vfvm_mesh_lattice_dev(int32_t n, uint64_t seed, double jitter, int32_t device, vfvm_mesh** out);
int vfvm_mesh_from_tets_dev(int64_t num_pts, int64_t num_nds, const double* pos,
                            int64_t num_tets, const int32_t* tets,
                            int32_t num_bnd, const int64_t* bnd_ptr, const int32_t* bnd_nid,
                            int32_t device, vfvm_mesh** out);

/* The vertices (i,j,k) of the n^3 jittered lattice above with lo <= (i,j,k) <= hi (a box of the
 * unit cube, at least 3 vertices per axis), Voronoi dual on GPU `device` (host if device < 0).
 * A coordinate on a face of the box is the unjittered lattice plane, so boxes that share a
 * face place its vertices identically; boundaries "x0".."z1" are the box's faces with the
 * ownership of vfvm_mesh_lattice.  lo = 0, hi = n - 1 gives vfvm_mesh_lattice's mesh.  Builds
 * the domains of a multi-domain lattice (Variables::new_obj + add_dom, src/base/vars.rs:50-69,
 * for box-shaped domains). */
int vfvm_mesh_lattice_box(int32_t n, uint64_t seed, double jitter, const int32_t* lo, const int32_t* hi,
                          int32_t device, vfvm_mesh** out);
/* Joins num_dom meshes (each with its dual) into one multi-domain mesh: domain d's vertices are
 * [dom_ptr[d], dom_ptr[d+1]) in the domain's own order (points, then nodes), so an unknown over
 * the joined vertices is the reference's per-domain unknown scalars with their blocks in domain
 * order (SteadyBase::initial_start); each domain keeps its own dual (block-diagonal neighbour
 * CSR).  itf: num_itf quadruples (dom_a, bnd_a, dom_b, bnd_b): boundary bnd_a of dom_a and bnd_b
 * of dom_b hold the same nodes (equal coordinates) and become interface itf (Interface::new,
 * src/base/geom_itf.rs:19-49; side a = dom_a, whose node order the interface keeps); the other
 * boundaries keep their order, domain by domain.  In the joined mesh num_pts = 0 and node ids
 * are vertex ids.  Flow operators and pressure-gradient coefficients are not supported on it. */
int vfvm_mesh_join(int32_t num_dom, const vfvm_mesh* const* doms, int32_t num_itf, const int32_t* itf,
                   vfvm_mesh** out);

int vfvm_mesh_info_get(const vfvm_mesh* mesh, vfvm_mesh_info* info);

/* Builds the pressure-gradient coefficients of the flow operators: for every vertex v,
 * int_V grad(p) dV = sum_m (c_x, c_y, c_z)_m p_m over the 2-ring m of v, with the skew
 * correction of the face values fully implicit (Tetrahedron::grad_coeff,
 * src/base/geom_tet.rs:295-347).  Needed before uploading a mesh for the flow operators. */
int vfvm_mesh_build_grad_coeff(vfvm_mesh* mesh);
/* The same coefficients built on GPU `device`, bitwise equal to the host build
 * (Tetrahedron::grad_coeff, src/base/geom_tet.rs:295-347). */
int vfvm_mesh_build_grad_coeff_dev(vfvm_mesh* mesh, int32_t device);

/* Array ids for vfvm_mesh_get (element type and length in brackets). */
enum vfvm_mesh_array {
    VFVM_MESH_POS = 0,       /* f64 [3N] AoS                         */
    VFVM_MESH_TETS = 1,      /* i32 [4T]                             */
    VFVM_MESH_BND_PTR = 2,   /* i64 [num_bnd+1]                      */
    VFVM_MESH_BND_NID = 3,   /* i32 [bnd_ptr[num_bnd]]               */
    VFVM_MESH_ROW_PTR = 4,   /* i32 [N+1]                            */
    VFVM_MESH_NGH = 5,       /* i32 [E]                              */
    VFVM_MESH_DIST = 6,      /* f64 [E]                              */
    VFVM_MESH_AREA = 7,      /* f64 [E]                              */
    VFVM_MESH_NORM = 8,      /* f64 [3E] SoA                         */
    VFVM_MESH_SKEW = 9,      /* f64 [3E] SoA                         */
    VFVM_MESH_GRAD_WGT = 10, /* f64 [3E] SoA                         */
    VFVM_MESH_VOL = 11,      /* f64 [N]                              */
    VFVM_MESH_CSKEW = 12,    /* f64 [3N] SoA                         */
    VFVM_MESH_BSF_PTR = 13,  /* i32 [N+1]                            */
    VFVM_MESH_BSF_VJ = 14,   /* i32 [S]                              */
    VFVM_MESH_BSF_VK = 15,   /* i32 [S]                              */
    VFVM_MESH_BSF_AREA = 16, /* f64 [S]                              */
    VFVM_MESH_BSF_NORM = 17, /* f64 [3S] SoA                         */
    VFVM_MESH_BSF_WGT = 18,  /* f64 [3S] SoA                         */
    VFVM_MESH_GC_PTR = 19,   /* i32 [N+1] pressure-gradient coefficients, CSR over vertices */
    VFVM_MESH_GC_VID = 20,   /* i32 [G] vertex m (the 2-ring of v, ascending)            */
    VFVM_MESH_GC_C = 21,     /* f64 [3G] SoA (c_x, c_y, c_z)                              */
    VFVM_MESH_DOM_PTR = 22,  /* i64 [num_dom+1] first vertex of each domain                  */
    VFVM_MESH_DOM_NPTS = 23, /* i64 [num_dom] points of each domain (its first vertices)     */
    VFVM_MESH_ITF_PTR = 24,  /* i64 [num_itf+1] interface node ranges                        */
    VFVM_MESH_ITF_VA = 25,   /* i32 [num_itf_nodes] vertex of the node's side-a copy          */
    VFVM_MESH_ITF_VB = 26,   /* i32 [num_itf_nodes] vertex of the node's side-b copy          */
    VFVM_MESH_ITF_DOM = 27   /* i32 [2 num_itf] domains of side a and side b                  */
};
int vfvm_mesh_get(const vfvm_mesh* mesh, int32_t array_id, void* dst);

/* CSV output of a field (write_scldom_csv / write_vecdom_csv / write_sclbnd_csv /
 * write_vecbnd_csv, src/base/write_csv.rs:5-100): file "<file_name>_<ts, 4 digits>.<file_type>",
 * header "x,y,z,value" (ncomp = 1) or "x,y,z,value_x,value_y,value_z" (ncomp = 3), one line per
 * vertex in vertex order (points, then nodes), or per node of boundary `bnd` (bnd >= 0, values
 * in the boundary's node order).  values: ncomp arrays of that length, back to back.
 * digits = 9 writes the reference's fixed 9-decimal format; digits = 0 writes every number
 * with 17 significant digits (reads back to the same double).  Host only. */
int vfvm_write_csv(const vfvm_mesh* mesh, int32_t bnd, int32_t ncomp, const double* values, const char* file_name,
                   const char* file_type, int32_t ts, int32_t digits);
/* The same field as an ASCII VTU file (write_{scl,vec}{dom,bnd}_vtu, src/base/write_vtu.rs):
 * a domain field on the mesh's tetrahedra (VTK type 10), a boundary field on one vertex cell
 * per node (type 1), point data "value" with ncomp components; numbers as vfvm_write_csv. */
int vfvm_write_vtu(const vfvm_mesh* mesh, int32_t bnd, int32_t ncomp, const double* values, const char* file_name,
                   const char* file_type, int32_t ts, int32_t digits);
void vfvm_mesh_destroy(vfvm_mesh* mesh);

/* ============================ device context ============================== */

typedef struct vfvm_ctx vfvm_ctx;

int vfvm_create(int32_t device, vfvm_ctx** out);
void vfvm_destroy(vfvm_ctx* ctx);

/* Uploads the dual geometry of `mesh` to HBM (SoA, CSR edge order) and resets
 * fields, operators and the pattern.  The mesh may be destroyed afterwards. */
int vfvm_mesh_upload(vfvm_ctx* ctx, const vfvm_mesh* mesh);

/* ---- fields (ScalarDomain / VectorDomain / ScalarBoundary / ScalarFlow) ---- */
enum vfvm_field_kind {
    VFVM_FIELD_CONSTANT = 0, /* SclDomType::Constant  (scl_dom.rs:9)               */
    VFVM_FIELD_FUNCTION = 1, /* SclDomType::Function  (values evaluated by caller) */
    VFVM_FIELD_UNKNOWN = 2   /* SclDomType::Unknown   (start assigned by driver)   */
};
/* values: N doubles (NULL -> all equal to `cnst`).  Gradients are computed on the
 * device by the least-squares gather of ScalarDomain::update_grad (scl_dom.rs:146). */
int vfvm_scldom_add(vfvm_ctx* ctx, int32_t kind, double cnst, const double* values, int32_t* id);
int vfvm_scldom_set(vfvm_ctx* ctx, int32_t id, const double* values);
int vfvm_scldom_get(vfvm_ctx* ctx, int32_t id, double* values);
/* Copies value into value_prev (ScalarDomain::update_prev, scl_dom.rs:141). */
int vfvm_scldom_update_prev(vfvm_ctx* ctx, int32_t id);
/* values: 3N doubles SoA (x[N], y[N], z[N]). */
int vfvm_vecdom_add(vfvm_ctx* ctx, int32_t kind, const double* cnst3, const double* values, int32_t* id);
int vfvm_vecdom_set(vfvm_ctx* ctx, int32_t id, const double* values);
int vfvm_vecdom_get(vfvm_ctx* ctx, int32_t id, double* values);
/* values: one double per node of boundary bnd_id, in the boundary's node order. */
int vfvm_sclbnd_add(vfvm_ctx* ctx, int32_t bnd_id, double cnst, const double* values, int32_t* id);
int vfvm_sclbnd_set(vfvm_ctx* ctx, int32_t id, const double* values);
/* Flow data (ScalarFlow, scl_flw.rs:3-22) with a prescribed velocity field:
 * the face mass flux of flow_helper::update_flux (flow_helper.rs:72-179). */
int vfvm_sclflw_add(vfvm_ctx* ctx, int32_t vel_id, int32_t pres_id, int32_t den_id,
                    int32_t visc_id, int32_t fce_id, int32_t* id);
/* Boundary velocity (VectorBoundary, src/base/vec_bnd.rs) for OpVecBndDirichlet: values
 * are 3 x count doubles SoA (x[count], y[count], z[count]) in the boundary's node order
 * (NULL -> every node takes cnst3[0..2]). */
int vfvm_vecbnd_add(vfvm_ctx* ctx, int32_t bnd_id, const double* cnst3, const double* values, int32_t* id);
/* New values of a boundary velocity (VectorBoundary::update_func, vec_bnd.rs), same layout. */
int vfvm_vecbnd_set(vfvm_ctx* ctx, int32_t id, const double* values);
/* flow_helper::update_flow (flow_helper.rs:3-179) for a flow with an unknown pressure: the
 * gradients of its fields, rc = V / a_P, the momentum term rc (grad p - fce_c) and the
 * Rhie-Chow face mass flux, from the current field values. */
int vfvm_update_flow(vfvm_ctx* ctx, int32_t flw_id);
/* Reads a flow's data: what = 0 face mass flux (E, CSR edge order), 1 rc (N),
 * 2 momentum term (3N SoA). */
int vfvm_sclflw_get(vfvm_ctx* ctx, int32_t flw_id, int32_t what, double* out);

/* ---- operators: one descriptor per reference operator struct ---- */
enum vfvm_op_kind {
    /* domain operators; geom_id = dom_id (0) */
    VFVM_OP_SCL_DOM_DIFFUSION = 1,        /* OpSclDomDiffusion{unk,diff,drv}      opscl_dom_diff.rs:7     a = diff, b = drv     */
    VFVM_OP_SCL_DOM_PRODUCTION = 2,       /* OpSclDomProduction{unk,prd}          opscl_dom_prod.rs        a = prd               */
    VFVM_OP_SCL_DOM_SOURCE = 3,           /* OpSclDomSource{unk,src}              opscl_dom_src.rs         a = src               */
    VFVM_OP_SCL_DOM_TIME = 4,             /* OpSclDomTime{unk,wgt}                opscl_dom_time.rs:7      a = wgt               */
    VFVM_OP_SCL_DOM_ADVECTION_UNITY = 5,  /* OpSclDomAdvectionUnity{unk,diff?,flw} opscl_dom_advunity.rs   a = diff or -1, b = flw */
    VFVM_OP_SCL_DOM_ADVECTION = 6,        /* OpSclDomAdvection{unk,wgt,diff,den,flw} opscl_dom_adv.rs:49  a = wgt, b = diff, c = flw (den = the flow's) */
    VFVM_OP_SCL_DOM_TIME_PREVIOUS = 7,    /* OpSclDomTimePrevious{unk,wgt}        opscl_dom_timeprev.rs    a = wgt (its previous values) */
    VFVM_OP_SCL_DOM_TIME_UNITY = 8,       /* OpSclDomTimeUnity{unk}               opscl_dom_timeunity.rs                         */
    /* boundary operators; geom_id = bnd_id */
    VFVM_OP_SCL_BND_DIRICHLET = 20,       /* OpSclBndDirichlet{unk,val}           opscl_bnd_dir.rs         a = sclbnd value      */
    VFVM_OP_SCL_BND_NEUMANN = 21,         /* OpSclBndNeumann{unk,flux}            opscl_bnd_neu.rs         a = sclbnd flux       */
    VFVM_OP_SCL_BND_DIFFUSION = 22,       /* OpSclBndDiffusion{unk,diff,drv}      opscl_bnd_diff.rs        a = diff, b = drv     */
    VFVM_OP_SCL_BND_PRODUCTION = 23,      /* OpSclBndProduction{unk,prd}          opscl_bnd_prod.rs        a = prd               */
    VFVM_OP_SCL_BND_SOURCE = 24,          /* OpSclBndSource{unk,src}              opscl_bnd_src.rs         a = src               */
    VFVM_OP_SCL_BND_TIME = 25,            /* OpSclBndTime{unk,wgt}                opscl_bnd_time.rs        a = wgt               */
    VFVM_OP_SCL_BND_ADVECTION_UNITY = 26, /* OpSclBndAdvectionUnity{unk,diff?,vel,flw} opscl_bnd_advunity.rs a = diff or -1, b = flw */
    VFVM_OP_SCL_BND_ADVECTION = 27,       /* OpSclBndAdvection{unk,wgt,diff,vel,den,flw} opscl_bnd_adv.rs a = wgt, b = diff, c = flw */
    VFVM_OP_SCL_BND_TIME_PREVIOUS = 28,   /* OpSclBndTimePrevious{unk,wgt}        opscl_bnd_timeprev.rs    a = wgt                */
    VFVM_OP_SCL_BND_TIME_UNITY = 29,      /* OpSclBndTimeUnity{unk}               opscl_bnd_timeunity.rs                         */
    /* flow operators (SteadyFlow, steady_flow.rs:133-240): unknown velocity vecdom `vel` and
     * pressure scldom `pres` of one flow (sclflw `flw`, whose den/visc/fce ids the operators
     * must name); domain operators geom_id = dom_id (0) */
    VFVM_OP_VEC_DOM_ADVECTION = 40,       /* OpVecDomAdvection{vel,den,visc,flw}  opvec_dom_adv.rs:47     unk = vel, a = den, b = visc, c = flw */
    VFVM_OP_VEC_DOM_GRADIENT = 41,        /* OpVecDomGradient{vel,pres}           opvec_dom_grad.rs:39    unk = vel, a = pres   */
    VFVM_OP_VEC_DOM_DIFFUSION = 42,       /* OpVecDomDiffusion{vel,visc}          opvec_dom_diff.rs:39    unk = vel, a = visc   */
    VFVM_OP_VEC_DOM_TRANSPOSE = 43,       /* OpVecDomTranspose{vel,visc}          opvec_dom_dtrn.rs:39    unk = vel, a = visc   */
    VFVM_OP_VEC_DOM_DILATATORY = 44,      /* OpVecDomDilatatory{vel,visc}         opvec_dom_ddil.rs:41    unk = vel, a = visc   */
    VFVM_OP_VEC_DOM_SOURCE = 45,          /* OpVecDomSource{vel,fce}              opvec_dom_src.rs:37     unk = vel, a = fce    */
    VFVM_OP_SCL_DOM_DIVERGENCE = 46,      /* OpSclDomDivergence{vel,pres,den}     opscl_dom_div.rs:68     unk = pres, a = vel, b = den */
    VFVM_OP_SCL_DOM_PRESSURE_STABILIZER = 47, /* OpSclDomPressureStabilizer{pres,den,fce,flw} opscl_dom_pstab.rs:52 unk = pres, a = den, b = fce, c = flw */
    VFVM_OP_SCL_DOM_PRESSURE_REFERENCE = 48,  /* OpSclDomPressureReference{pres,nid,pres_ref} opscl_dom_pref.rs:48 unk = pres, a = nid, value = pres_ref */
    VFVM_OP_VEC_DOM_TIME = 49,            /* OpVecDomTime{vel,den}                opvec_dom_time.rs:37    unk = vel, a = den; first
                                             in the list: a TransientFlow system, whose divergence operators also
                                             apply their time part (opscl_dom_div.rs:40-66, opscl_bnd_div.rs:40-70) */
    /* flow boundary operators; geom_id = bnd_id */
    VFVM_OP_VEC_BND_DIRICHLET = 60,       /* OpVecBndDirichlet{vel,val}           opvec_bnd_dir.rs:56     unk = vel, a = vecbnd value */
    VFVM_OP_SCL_BND_DIVERGENCE = 61,      /* OpSclBndDivergence{vel,pres,den}     opscl_bnd_div.rs:73     unk = pres, a = vel, b = den */
    VFVM_OP_SCL_BND_PRESSURE_STABILIZER = 62, /* OpSclBndPressureStabilizer{pres,den,fce,flw} opscl_bnd_pstab.rs:49 unk = pres, a = den, b = fce, c = flw */
    /* scalar interface operators on a joined mesh (vfvm_mesh_join); geom_id = itf.  The fields
     * are joined-mesh fields: `unk` holds the reference's (unk_a, unk_b) pair (its values on
     * side a's and side b's domains), `a` the (coef_a, coef_b) pair.  The continuity operator
     * fills the rows of side a's node copies, every other one the rows of side b's copies (the
     * combined control volume); no domain or boundary operator touches an interface node. */
    VFVM_OP_SCL_ITF_CONTINUITY = 70,      /* OpSclItfContinuity{itf,unk_a,unk_b}   opscl_itf_cont.rs:38-58                         */
    VFVM_OP_SCL_ITF_DIFFUSION = 71,       /* OpSclItfDiffusion{itf,unk,diff,drv}   opscl_itf_diff.rs:49-113  a = diff, b = drv (= unk) */
    VFVM_OP_SCL_ITF_PRODUCTION = 72,      /* OpSclItfProduction{itf,unk,prd}       opscl_itf_prod.rs:46-96   a = prd               */
    VFVM_OP_SCL_ITF_SOURCE = 73,          /* OpSclItfSource{itf,unk,src}           opscl_itf_src.rs:44-87    a = src               */
    VFVM_OP_SCL_ITF_TIME = 74,            /* OpSclItfTime{itf,unk,wgt}             opscl_itf_time.rs:42-80   a = wgt (apply_time)  */
    VFVM_OP_SCL_ITF_TIME_PREVIOUS = 75,   /* OpSclItfTimePrevious{itf,unk,wgt}     opscl_itf_timeprev.rs:42-82 a = wgt (its previous values) */
    VFVM_OP_SCL_ITF_TIME_UNITY = 76,      /* OpSclItfTimeUnity{itf,unk}            opscl_itf_timeunity.rs:35-73                    */
    VFVM_OP_SCL_ITF_ADVECTION = 77,       /* OpSclItfAdvection{itf,unk,wgt,diff,den,flw} opscl_itf_adv.rs:61-150 a = wgt, b = diff, c = flw
                                             (den = the flow's), value = flowing side (0: a, 1: b) */
    VFVM_OP_SCL_ITF_ADVECTION_UNITY = 78  /* OpSclItfAdvectionUnity{itf,unk,diff?,flw} opscl_itf_advunity.rs:59-137 a = diff or -1,
                                             b = flw, value = flowing side (0: a, 1: b) */
};
typedef struct vfvm_op {
    int32_t kind;     /* vfvm_op_kind                                         */
    int32_t geom_id;  /* dom_id (0) or bnd_id                                 */
    int32_t unk_id;   /* unknown scalar (scldom id) or velocity (vecdom id)   */
    int32_t a, b, c;  /* per-kind field ids, see vfvm_op_kind (-1: unused)    */
    double value;     /* per-kind constant (reference pressure), else 0      */
} vfvm_op;
/* An operator list that holds flow operators describes the coupled velocity-pressure system
 * of SteadyFlow: unknown vector [u (N), v (N), w (N), p (N)] (steady_flow.rs:94-104), the
 * eight domain operators once each in the order of assemble_matrix (steady_flow.rs:244-254)
 * (TransientFlow: VEC_DOM_TIME before them, transient_flow.rs:250-260, solved by
 * vfvm_transient_solve),
 * for every boundary the triple {VEC_BND_DIRICHLET, SCL_BND_DIVERGENCE,
 * SCL_BND_PRESSURE_STABILIZER} (steady_flow.rs:256-260), and at most one
 * SCL_DOM_PRESSURE_REFERENCE (applied last, steady_flow.rs:297-300).  The mesh needs its
 * pressure-gradient coefficients (vfvm_mesh_build_grad_coeff).  The velocity rows of the
 * points hold [c N + ({v} u ngh(v)) | 3N + grad_coeff(v)], the pressure rows
 * [c N + grad_coeff(v), c = 0..2 | 3N + 3-ring(v)], all ascending; every entry is the sum of
 * its pushes in the reference's push order.  Scalar operators in the same list (the heat or
 * species part of SteadyHeatFlow / TransientHeatMassFlow, steady_heatflow.rs:468-505) act
 * on further scalar unknowns, each a block of N rows after the pressure, in order of first
 * appearance: x = [u, v, w, p, s_1, s_2, ...] (at most two, block Jacobi's b <= 6). */

/* Fixes the unknown's block start (SteadyBase::initial_start), builds the row-sorted
 * CSR pattern of the operator list and the edge->slot / diag->slot maps on the host,
 * uploads them (LinearBase::update_sparse).  The list is applied in array order, as
 * assemble_matrix applies its operators. */
int vfvm_pattern_build(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, int64_t* nnz);
int vfvm_pattern_get(vfvm_ctx* ctx, int32_t* row_ptr, int32_t* col);
/* Size of the built system: n = mat_size (rows), nnz (LinearBase::update_sparse's pattern). */
int vfvm_system_info(vfvm_ctx* ctx, int64_t* n, int64_t* nnz);

/* update_grad + update_flow + assemble_matrix + build_matrix on the device: writes
 * the CSR values and b in place, in a fixed per-row summation order, no atomics.
 * dt > 0 enables the apply_time operators (dt <= 0 with a TIME op is an error). */
int vfvm_assemble(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, double dt);
int vfvm_matrix_get(vfvm_ctx* ctx, double* vals, double* b);

/* ---- linear solve (LinearGmres) ---- */
enum vfvm_precision { VFVM_FP64 = 0, VFVM_FP32_REFINED = 1 };
/* ILU0: ILU(0) of the matrix in multicolour order (greedy colouring of the row graph,
 * rows sorted by (colour, index)); its triangular sweeps run one launch per colour. */
/* BJACOBI: block Jacobi over the unknowns of one vertex (b = mat_size / N unknowns, entries
 * v, N + v, ..., (b-1) N + v; b <= 6): the b x b diagonal blocks inverted at set-up (Gauss-Jordan
 * with partial pivoting; a singular block is replaced by the identity). */
enum vfvm_precond { VFVM_PRECOND_NONE = 0, VFVM_PRECOND_JACOBI = 1, VFVM_PRECOND_ILU0 = 2, VFVM_PRECOND_BJACOBI = 3 };
/* AUTO: CGS2 (one re-orthogonalisation) for FP64 GMRES, CGS for the FP32 inner solve. */
enum vfvm_ortho { VFVM_ORTHO_AUTO = 0, VFVM_ORTHO_CGS = 1, VFVM_ORTHO_CGS2 = 2 };
typedef struct vfvm_solve_opts {
    double tol;          /* relative residual ||b-Ax||/||b|| target (default 1e-6, linear_gmres.rs:21) */
    int32_t restart;     /* Krylov dimension per cycle (default 50)                    */
    int32_t max_outer;   /* restart cycles (default 200)                               */
    int32_t precision;   /* vfvm_precision                                             */
    int32_t precond;     /* vfvm_precond (default Jacobi, linear_gmres.rs:79)          */
    double inner_tol;    /* FP32 inner relative tolerance per refinement step (1e-5)   */
    int32_t max_refine;  /* maximum refinement steps (FP32_REFINED)                    */
    int32_t ortho;       /* vfvm_ortho: Gram-Schmidt variant (0 = auto)                */
} vfvm_solve_opts;
typedef struct vfvm_solve_stats {
    int32_t iterations;      /* Arnoldi steps                               */
    int32_t restarts;        /* restart cycles                              */
    int32_t refinements;     /* outer FP64 refinement steps                 */
    int32_t status;          /* vfvm_status of the solve                    */
    double rel_residual;     /* final ||b-Ax||/||b|| computed in FP64       */
    double seconds;          /* device wall time of the solve (HIP events) */
    double setup_seconds;    /* preconditioner set-up part of `seconds`     */
} vfvm_solve_stats;
void vfvm_solve_opts_default(vfvm_solve_opts* opts);
/* Solves the assembled system; x: host array, initial guess in, solution out
 * (LinearBase::solve_matrix, linear_gmres.rs:76-84). */
int vfvm_solve(vfvm_ctx* ctx, const vfvm_solve_opts* opts, double* x, vfvm_solve_stats* stats);
/* Same, on a device-resident x (hipMalloc'd, mat_size doubles) on the context stream. */
int vfvm_solve_device(vfvm_ctx* ctx, const vfvm_solve_opts* opts, double* x_dev, vfvm_solve_stats* stats);
/* No reference counterpart.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): vfvm_solve_device with
 * no allocation of any kind.  The Krylov workspace lives in a fixed array of 128 MiB per GPU
 * (held from the library's load until the process ends); a system needs about
 * 8 (n (restart + 5) + nnz) bytes of it (nnz only with Jacobi) plus a few MB of partial sums,
 * and larger ones are rejected with VFVM_ERR_INVALID_ARGUMENT.  FP64 with the Jacobi
 * preconditioner or none; launches are eager (no graphs), stats.seconds is host wall time.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two threads,
 * two contexts) race and must be serialised by the caller. */
int vfvm_solve_device_noalloc(vfvm_ctx* ctx, const vfvm_solve_opts* opts, double* x_dev, vfvm_solve_stats* stats);
/* r = b - A x in FP64 for a host x; returns ||r||_2 / ||b||_2. */
int vfvm_residual(vfvm_ctx* ctx, const double* x, double* rel_residual);
/* z = M^-1 r for the assembled matrix with the preconditioner of opts (precond, and
 * precision: FP32_REFINED factors and applies ILU(0) in FP32), host arrays.  No reference
 * counterpart (the reference's Jacobi lives inside faer_gmres); exposed for parity tests. */
int vfvm_precond_apply(vfvm_ctx* ctx, const vfvm_solve_opts* opts, const double* r, double* z);
/* The ILU(0) row order of the built pattern: perm[t] = row of A at position t, color_ptr[ncolors
 * + 1] = first position of each colour.  Multicolour order: rows sorted by (colour, Morton code),
 * no two rows of one colour coupled.  Tiled order (scalar systems on a mesh): rows sorted by
 * (tile colour, tile, level, Morton code); the "colours" are the tile colours, whose rows couple
 * only inside one tile (vfvm_ilu_tiles).  For the parity tests; no reference counterpart. */
int vfvm_ilu_order(vfvm_ctx* ctx, int32_t* perm, int32_t* color_ptr, int32_t* ncolors);
/* The tiles of the tiled ILU(0) order (ntiles = 0 for the multicolour order): tile k holds the
 * positions [tile_ptr[k], tile_ptr[k+1]) (tile_ptr: ntiles + 1 entries, at most n + 1), level[t]
 * = level of position t inside its tile (n entries); rows of one tile and one level are not
 * coupled.  For the parity tests; no reference counterpart. */
int vfvm_ilu_tiles(vfvm_ctx* ctx, int32_t* tile_ptr, int32_t* ntiles, int32_t* level);
/* The one-launch ILU(0) sweeps of the tiled order wait for the completion flags of adjacent
 * tiles with at most `polls` polls each (2^17 by default; polls < 0: every wait for another tile
 * gives up at once).  A sweep whose wait gives up writes NaN, so the solve stops with
 * VFVM_ERR_LINEAR_NON_FINITE instead of hanging; negative limits exist to test that path.
 * *lapsed (may be null) receives the number of waits that gave up so far.  No reference
 * counterpart. */
int vfvm_ilu_wait_limit(vfvm_ctx* ctx, int32_t polls, uint64_t* lapsed);

/* ---- caller-assembled systems (LinearBase, src/linear/linear_base.rs:5-10) ----
 * For a solver whose operators are not on the device: it hands its triplet list, right-hand
 * side and initial guess to the device GMRES, as the reference's drivers hand theirs to
 * LinearGmres.  The system replaces the context's operator-assembled one (vfvm_assemble then
 * fails until vfvm_pattern_build is called again); vfvm_solve, vfvm_residual and
 * vfvm_precond_apply work on it.  No mesh is needed (block Jacobi needs one: its blocks are the
 * unknowns of a vertex). */
/* LinearBase::update_sparse (linear_gmres.rs:53-64): the pattern of the triplets (rows[k],
 * cols[k]), k < ntrip, of a mat_size x mat_size matrix (0-based, as usize in the reference),
 * duplicates merged, rows sorted by column; nnz: the number of distinct entries. */
int vfvm_linear_update_sparse(vfvm_ctx* ctx, int64_t mat_size, int64_t ntrip, const int64_t* rows, const int64_t* cols,
                              int64_t* nnz);
/* LinearBase::build_matrix (linear_gmres.rs:66-74): the matrix of a triplet list with the
 * (row, col) sequence given to update_sparse (its argsort is reused, as new_from_argsort
 * does): vals[k] of triplet k, duplicates summed in list order.  Host array; the _device
 * variant reads a device array of ntrip doubles. */
int vfvm_linear_build_matrix(vfvm_ctx* ctx, int64_t ntrip, const double* vals);
int vfvm_linear_build_matrix_device(vfvm_ctx* ctx, int64_t ntrip, const double* vals_dev);
/* LinearBase::solve_matrix (linear_gmres.rs:76-84): solves A x = b from x_init (NULL: zero)
 * with the options' GMRES; host arrays of mat_size doubles.  The _device variant reads b_dev
 * and takes x_dev as initial guess in, solution out (device arrays). */
int vfvm_linear_solve_matrix(vfvm_ctx* ctx, const vfvm_solve_opts* opts, const double* b, const double* x_init, double* x,
                             vfvm_solve_stats* stats);
int vfvm_linear_solve_matrix_device(vfvm_ctx* ctx, const vfvm_solve_opts* opts, const double* b_dev, double* x_dev,
                                    vfvm_solve_stats* stats);

/* ---- drivers ---- */
/* Host closure of function fields that depend on unknowns or time (Variables::_update_func,
 * steady_base.rs:242-261): called with the time t before every assembly of the drivers
 * (vfvm_steady_solve, vfvm_transient_solve, vfvm_transient_adaptive), after the unknowns
 * hold the current iterate; it may read fields (vfvm_scldom_get ...) and set function fields
 * (vfvm_scldom_set, vfvm_vecdom_set).  NULL removes it. */
typedef void (*vfvm_update_hook)(void* user, double t);
int vfvm_set_update_hook(vfvm_ctx* ctx, vfvm_update_hook hook, void* user);
/* Convergence scaling (src/base/scaling.rs): a reference scale of an unknown scalar or vector
 * field, a lower bound of the field's scale in the scaled Picard residual (scaling.rs:88-111)
 * and the adaptive error norm (Variables::set_scldom_scale / set_vecdom_scale, vars.rs:203-213;
 * VFVM_ERR_VARIABLES_NON_POSITIVE_SCALE for scale <= 0 or NaN).  For a pressure it is the
 * reference pressure difference. */
int vfvm_set_scldom_scale(vfvm_ctx* ctx, int32_t id, double scale);
int vfvm_set_vecdom_scale(vfvm_ctx* ctx, int32_t id, double scale);
/* Variables::scale_off (vars.rs:31): off != 0 selects the unscaled Picard residual
 * |x_new - x_old|_2 / (|x_old|_2 + 1e-10) over the whole system vector (scaling.rs:116-122) and
 * the per-block error norm of the adaptive driver (transient_adaptive.rs:257-270). */
int vfvm_set_scale_off(vfvm_ctx* ctx, int32_t off);

typedef struct vfvm_driver_stats {
    int32_t picard_iterations;   /* total Picard iterations                   */
    int32_t linear_iterations;   /* total Arnoldi steps                       */
    double residual;             /* last scaled residual (scaling.rs:113)     */
    double time_assembly;        /* "Matrix Assembly" bucket, seconds         */
    double time_solve;           /* "Matrix Solve" bucket, seconds            */
} vfvm_driver_stats;
/* SteadyBase::solve (steady_base.rs:16-135) for the unknown of the built pattern:
 * Dirichlet initial values, assemble, Picard loop with damping, scaled residual. */
int vfvm_steady_solve(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts,
                      int32_t max_iter, double tol, double damp, vfvm_driver_stats* stats);
/* TransientBase::solve (transient_base.rs:16-160): implicit Euler, Picard per step. */
int vfvm_transient_solve(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts,
                         double dt, int32_t num_ts, int32_t max_iter, double tol, double damp,
                         vfvm_driver_stats* stats);

/* ---- adaptive time stepping (TransientAdaptive, src/transient/transient_adaptive.rs) ----
 * Backward Euler with the operator list's time operators, the local truncation error from
 * the linear-extrapolation predictor through the last two accepted states
 * (LTE ~ dt / (2 dt + dt_prev) |y - y_p|; constant predictor and factor 1/2 on the first
 * step), error norm max over the unknown fields of max_i |LTE_i| / (atol + rtol s_f) with the
 * field scales of scaling.rs, PI controller fac = safety err^-0.35 err_prev^0.2 in
 * [fac_min, fac_max], rejection on err > 1 (fac = max(fac_min, safety err^-1/2)) or on
 * Picard failure (dt fac_fail), steps shortened to land on the output times.  Works for a
 * scalar transient list and for a TransientFlow list. */
typedef struct vfvm_adaptive_opts {
    double t_end, dt_init, dt_min, dt_max;  /* AdaptiveOptions (transient_adaptive.rs:42-60) */
    double rtol, atol;
    double safety, fac_min, fac_max, fac_fail;
    int32_t max_reject;      /* consecutive rejections before giving up                 */
    int32_t max_iter;        /* Picard                                                  */
    double tol, damp;        /* Picard                                                  */
    int32_t fixed;           /* 1: dt = dt_init throughout (output times still hit)     */
    int32_t num_out;         /* output times in out_times (-1: 11 evenly spaced to t_end) */
    const double* out_times;
    int32_t num_exclude;     /* scldom ids left out of the error norm (e.g. the pressure) */
    int32_t reserved;
    const int32_t* scl_exclude;
} vfvm_adaptive_opts;
typedef struct vfvm_adaptive_stats {
    int32_t accepted, rejected_error, rejected_picard, picard_iterations;
    int32_t linear_iterations, outputs;
    double t_final, time_assembly, time_solve;
} vfvm_adaptive_stats;
/* Called after each accepted step with the state at t (value_prev still holds t - dt). */
typedef void (*vfvm_step_hook)(void* user, double t, double dt);
/* AdaptiveOptions::new (transient_adaptive.rs:62-73). */
void vfvm_adaptive_opts_default(double dt_init, double t_end, vfvm_adaptive_opts* opts);
/* TransientAdaptive::solve_adaptive_with (transient_adaptive.rs:81-350).  steps_out (may be
 * NULL): (t, dt, err) of the first max_steps accepted steps.  lin: the linear solves. */
int vfvm_transient_adaptive(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* lin,
                            const vfvm_adaptive_opts* opts, vfvm_step_hook hook, void* user,
                            double* steps_out, int32_t max_steps, vfvm_adaptive_stats* stats);

/* ---- discrete adjoint (no reference counterpart: the reference has no adjoint) ----
 * Forward map of a linear problem (no operator coefficient depends on the unknown):
 *   steady    (num_ts = 0):  A T = f
 *   transient (num_ts > 0):  A T^n = B T^(n-1) + f,  n = 1..num_ts  (implicit Euler,
 *                            TransientBase::solve, transient_base.rs:16-160, one linear
 *                            solve per step, initial guess T^(n-1))
 * with A, f from the operator list (as vfvm_assemble) and B = diag(-w V / dt) from its TIME
 * operators.  For an objective J whose derivative `seed` = dJ/dT^num_ts is given, the
 * backward pass solves A^T l^N = seed, A^T l^n = B^T l^(n+1) (transposed GMRES, initial
 * guess l^(n+1)), replaying each checkpoint segment forward to recover T^n, and returns
 *   dJ/dp = sum_n  l^n . (df/dp + dB/dp T^(n-1) - dA/dp T^n)
 * for the requested parameters, accumulated per directed edge and per vertex on the device
 * with no atomics.  Supported operators: DIFFUSION, ADVECTION_UNITY, SOURCE, TIME (dom and
 * bnd), DIRICHLET, NEUMANN; PRODUCTION (whose right-hand side reads the unknown's gradient)
 * and coefficients equal to the unknown are rejected with VFVM_ERR_OPERATOR_UNSUPPORTED. */
typedef struct vfvm_adjoint_stats {
    int32_t steps;                 /* time steps of the forward map (0: steady)            */
    int32_t checkpoints;           /* states kept by the forward pass                      */
    int32_t forward_iterations;    /* Arnoldi steps of the forward pass                    */
    int32_t recompute_iterations;  /* Arnoldi steps of the segment replays                 */
    int32_t adjoint_iterations;    /* Arnoldi steps of the transposed solves               */
    int32_t recomputed_steps;      /* forward steps replayed by the backward pass          */
    int32_t replay_mismatches;     /* entries of replayed states that differ from the
                                      stored checkpoint they must reproduce (expected 0:
                                      the kernels are deterministic)                      */
    int32_t reserved;
    double forward_seconds;        /* time-step loop of vfvm_adjoint_forward (wall)        */
    double backward_seconds;       /* vfvm_adjoint_backward after its set-up (wall)        */
    double recompute_seconds;      /* replay part of backward_seconds                      */
    double adjoint_solve_seconds;  /* transposed-solve part of backward_seconds            */
    double setup_seconds;          /* host pattern/colouring set-up of both passes         */
} vfvm_adjoint_stats;
/* Runs the forward map from the unknown's current values (Dirichlet values applied first,
 * as vfvm_transient_solve), keeps T^0, T^K, T^2K, ... (K = checkpoint_every >= 1) and the
 * final state, and leaves T^num_ts in the unknown. */
int vfvm_adjoint_forward(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts,
                         double dt, int32_t num_ts, int32_t checkpoint_every, vfvm_adjoint_stats* stats);
enum vfvm_wrt_kind {
    VFVM_WRT_SCLDOM = 0,   /* per-vertex values of a scalar field (N entries): diffusion
                              coefficient, source or time weight, every use summed        */
    VFVM_WRT_SCLBND = 1,   /* per-node Dirichlet values of a boundary field (boundary size) */
    VFVM_WRT_INITIAL = 2   /* initial state T^0 as given before the Dirichlet values (N)    */
};
typedef struct vfvm_adjoint_wrt {
    int32_t kind;  /* vfvm_wrt_kind                                  */
    int32_t id;    /* field id (ignored for VFVM_WRT_INITIAL)        */
    double* grad;  /* output, host memory                            */
} vfvm_adjoint_wrt;
/* Backward pass of the last vfvm_adjoint_forward.  seed: host array of N entries.  opts:
 * the transposed solves' options.  The unknown holds T^num_ts again on return. */
int vfvm_adjoint_backward(vfvm_ctx* ctx, const vfvm_solve_opts* opts, const double* seed, int32_t nwrt,
                          const vfvm_adjoint_wrt* wrt, vfvm_adjoint_stats* stats);

/* ---- affine function fields ----
 * Variables::add_vecdom_func (src/base/vars.rs:153) for a closure affine in one scalar field:
 * vec := f0 + df * scl per component (f0, df: 3N SoA host arrays), evaluated on the device
 * where the drivers evaluate function fields (_update_func, before every assembly; before
 * the host update hook when both are set) and once at registration.  The vector field must
 * be a FUNCTION field.  f0 = df = NULL removes the closure.  The Boussinesq body force of a
 * natural-convection list is fce = (0, 0, gamma (T - T_ref)): f0 = (0, 0, -gamma T_ref),
 * df = (0, 0, gamma). */
int vfvm_vecdom_func_affine(vfvm_ctx* ctx, int32_t vec_id, int32_t scl_id, const double* f0, const double* df);

/* ---- adjoint of the coupled flow system (no reference counterpart) ----
 * Forward map: the Picard-converged coupled flow list (SteadyFlow / TransientFlow /
 * TransientHeat(Mass)Flow operator lists, as vfvm_steady_solve / vfvm_transient_solve run
 * them), i.e. the states x^n = [u, v, w, p, s_1, s_2] with
 *     R(x^n; x^(n-1), theta) = A(c(x^n)) x^n - b(c(x^n), x^(n-1), theta) = 0,  n = 1..num_ts
 * (num_ts = 0: one steady state), where c(x) are the state-dependent coefficients: the
 * Rhie-Chow face mass flux, rc = V / a_P, the velocity gradients and the body force (an
 * affine closure of vfvm_vecdom_func_affine; a body force that a host hook sets is taken as
 * a fixed field).  Objective J = sum_n w_n seed . x^n (steady: seed . x).
 * Backward: J(x)^T l^n = w_n seed - D l^(n+1) with the full Jacobian J = A + E (E: the
 * derivative through c, applied by vector-Jacobian kernels of the flux, rc, the pressure
 * stabiliser, the advection coefficients, the transpose / dilatatory / source right-hand
 * sides and the least-squares gradients) and D = dR^(n+1)/dx^n (the time operators'
 * rho V / dt, w V / dt), solved by GCR on J^T preconditioned with the transposed GMRES on
 * A^T; each checkpoint segment is replayed to recover x^n.  Coefficient fields must not
 * depend on the unknowns (constant or function fields); a PRODUCTION operator with a
 * non-zero coefficient is rejected. */
/* Called after each adjoint state with its step (0: steady), the outer iterations and inner
 * Arnoldi steps it took and its final ||s - J^T l|| / ||s||. */
typedef void (*vfvm_adjoint_hook)(void* user, int32_t step, int32_t outer, int32_t inner_steps, double rel_residual);
typedef struct vfvm_flow_adjoint_opts {
    vfvm_solve_opts inner;  /* transposed solves on A^T (the outer preconditioner)       */
    double tol;             /* outer: ||s - J^T l|| <= tol ||s|| per state               */
    int32_t max_iter;       /* outer iterations per state                                */
    int32_t restart;        /* outer directions kept before a restart (1..64)            */
    vfvm_adjoint_hook on_state;  /* may be NULL                                          */
    void* user;
} vfvm_flow_adjoint_opts;
typedef struct vfvm_flow_adjoint_stats {
    vfvm_adjoint_stats base;          /* steps, checkpoints, Arnoldi steps and times as the
                                         linear adjoint reports them (adjoint_iterations:
                                         the inner transposed solves)                     */
    int32_t picard_iterations;        /* forward pass                                     */
    int32_t recompute_picard_iterations;
    int32_t outer_iterations;         /* GCR iterations over every state                  */
    int32_t reserved;
    double objective;                 /* J of the forward pass                            */
    double max_outer_residual;        /* largest final ||s - J^T l|| / ||s|| of the states */
} vfvm_flow_adjoint_stats;
/* Runs the driver of the flow list from the fields' current values (as vfvm_steady_solve
 * for num_ts = 0, vfvm_transient_solve otherwise: Picard max_iter / tol / damp per step),
 * keeps x^0, x^K, x^2K, ... (K = checkpoint_every), and evaluates J: seed (n entries, may be
 * NULL: J = 0), weights (num_ts entries, w_n of step n; NULL: all 1; ignored when steady). */
int vfvm_flow_adjoint_forward(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, const vfvm_solve_opts* opts, double dt,
                              int32_t num_ts, int32_t max_iter, double tol, double damp, int32_t checkpoint_every,
                              const double* seed, const double* weights, vfvm_flow_adjoint_stats* stats);
/* Backward pass of the last vfvm_flow_adjoint_forward with the same seed / weights.  wrt:
 * VFVM_WRT_SCLBND (Dirichlet values of a coupled scalar's boundary field) and
 * VFVM_WRT_VECBND (velocity Dirichlet values, 3 x nodes SoA).  The fields hold the final
 * state again on return. */
enum { VFVM_WRT_VECBND = 3 };
int vfvm_flow_adjoint_backward(vfvm_ctx* ctx, const vfvm_flow_adjoint_opts* opts, const double* seed, const double* weights,
                               int32_t nwrt, const vfvm_adjoint_wrt* wrt, vfvm_flow_adjoint_stats* stats);
/* The residual r = A(c(x)) x - b(c(x), x_prev) of the flow list's system at the state x
 * (host arrays of n entries; x_prev: previous state of a transient list, NULL when dt = 0),
 * and y = J(x)^T l (J = A + E as above).  The pattern must have been built for the list; the
 * fields hold x on return. */
int vfvm_flow_residual_eval(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, double dt, const double* x, const double* x_prev,
                            double* r);
int vfvm_flow_jt_apply(vfvm_ctx* ctx, const vfvm_op* ops, int32_t nops, double dt, const double* x, const double* x_prev,
                       const double* lam, double* y);

/* ---- measurement ---- */
/* Per-launch HIP-event timing on the context stream: bit k of `mask` times kernel class k
 * (-1: all classes, 0: off).  Each timed launch is bracketed by two events. */
/* SPMV: the Krylov SpMV kernel; MDOT: the multi-dot of the Gram-Schmidt pass; UPDATE: the
 * w -= V h update; ASSEMBLE: the row-gather assembly; PRECOND: one ILU(0) triangular-sweep launch. */
enum vfvm_kernel_class { VFVM_K_SPMV = 0, VFVM_K_MDOT = 1, VFVM_K_ASSEMBLE = 2, VFVM_K_PRECOND = 3, VFVM_K_UPDATE = 4, VFVM_K_COUNT = 5 };
int vfvm_timing_enable(vfvm_ctx* ctx, int32_t mask);
/* Returns launches and total milliseconds of one kernel class since the last reset, and
 * the algorithmic bytes those launches moved (per the byte model of each kernel). */
int vfvm_timing_get(vfvm_ctx* ctx, int32_t kclass, int64_t* launches, double* total_ms, double* bytes);
int vfvm_timing_reset(vfvm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* VFVM_H */
