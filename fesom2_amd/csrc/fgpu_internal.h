// Internal declarations shared by the translation units of libfesom_gpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fesom_gpu.h"

#define FGPU_R_EARTH 6367500.0     // oce_modules.F90:29
#define FGPU_EPS_AB 0.1            // oce_modules.F90:105
#define FGPU_VCPW 4.2e6            // oce_modules.F90:31
#define FGPU_MAX_NL 128            // columns deeper than this are rejected
#define NXCD 8                     // XCDs of an MI355X; consecutive blocks alternate over them

// One block per unit (column or edge) over a grid padded to a multiple of NXCD:
// blocks b, b + NXCD, ... run on one XCD and take one contiguous range of
// units, so neighbouring units share that XCD's L2. Units at or past the real
// count are the padding and exit.
__device__ __forceinline__ int xcd_unit()
{
    const int per = (int)(gridDim.x / NXCD);
    return (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
}
static inline unsigned xcd_grid(int units) { return (unsigned)((units + NXCD - 1) / NXCD * NXCD); }

// tracer scheme codes
enum { HOR_MFCT = 0, HOR_UPW1 = 1, HOR_MUSCL = 2 };
enum { VER_QR4C = 0, VER_UPW1 = 1, VER_CDIFF = 2, VER_PPM = 3 };
enum { LIM_FCT = 0, LIM_NON = 1 };

// Mesh as seen by the kernels: 0-based connectivity, level numbers 1-based.
struct DevMesh {
    int N2, F2, E2, L, NL, kmax;
    const int *elem_nodes;   // F2*3
    const int *edges;        // E2*2
    const int *edge_tri;     // E2*2, -1 = none
    const int *einfo;        // E2*4: nl1, nu1, nl2, nu2 (reference names, oce_adv_tra_hor.F90:104-128)
    const int *erange;       // E2*2: nu12, nl12 of the max/min rule (oce_adv_tra_driver.F90:154-156)
    const int *updn;         // E2*2 up/downwind triangles, -1 = none
    const int *nie;          // N2*kmax nod_in_elem2D
    const int *nie_num;      // N2
    const int *nlev, *ulev;  // F2
    const int *nlevn, *ulevn, *nlevn_min, *ulevn_max;  // N2
    const int *ne_ptr;       // N2+1 node->edge gather lists
    const int *ne_idx;       // 2*e + side (side 0: node is edges(1,e)), edges ascending
    const int *ne_rec;       // 8 per ne_idx entry: q, other node, nu12, nl12, nl1, nu1, nl2, nu2
    const int *node_rec;     // 4 per node: ne_ptr, degree, ulevels_nod2D, nlevels_nod2D
    const int *nblay;        // N2: nboundary_lay of MUSCL (oce_muscl_adv.F90:91-157)
    const int *fct_pad;      // N2*2: levels k < fct_pad[2n] or k >= fct_pad[2n+1] have an adjacent element
                             // without level k, whose FCT bounds are the +-1e3 padding (oce_adv_tra_fct.F90:148-178)
    const double *elem_area, *elem_cos, *edge_dxdy, *ecd, *gsca;
    const double *area, *areasvol;        // N2*NL
    const double *hnode, *hnode_new;      // N2*L
    const double *helem;                  // F2*L
    const double *zbar3d;                 // N2*NL
    const double *Z3d;                    // N2*L
    const double *zbar_n_bot;             // N2
    const double *zbar_e_bot, *zbar_e_srf;  // F2
    // packed per-edge records of the fused edge pass (one scalar batch per edge)
    const int *erec_i;       // E2*ERI
    const double *erec_d;    // E2*ERD
    const double *zn_vd;     // N2*L: Z levels rebuilt from hnode_new for the implicit diffusion
    const double *zbn_vd;    // N2*NL: the interface depths of the same rebuild (Redi K33 weights)
    // processing order of the fused step's column blocks (a permutation of the
    // nodes / edges; results do not depend on it): along a Hilbert curve over
    // the coordinates, so that the columns running at once on one XCD are
    // spatial neighbours and share their neighbour gathers in its L2
    const int *node_ord;     // N2
    const int *edge_ord;     // E2
};
// erec_i fields
enum { ER_N1, ER_N2, ER_EL1, ER_EL2, ER_NL1, ER_NU1, ER_NL2, ER_NU2, ER_UA0, ER_UA1, ER_UA2, ER_DA0, ER_DA1,
       ER_DA2, ER_RULE, ER_NZMIN, ER_NZMAX, ER_U1, ER_U2, ER_NL1N, ER_NL2N, ER_NBL1, ER_NBL2, ER_I_PAD3, ERI };
// erec_d fields
enum { ER_DX1, ER_DY1, ER_DX2, ER_DY2, ER_A, ER_EDX, ER_EDY, ER_GU0, ER_GU1, ER_GU2, ER_GU3, ER_GU4, ER_GU5,
       ER_GD0, ER_GD1, ER_GD2, ER_GD3, ER_GD4, ER_GD5, ER_D_PAD, ERD };

struct DevDyn {
    const double *uv, *w, *w_e, *w_i, *Kv;
    const double *heat_flux, *water_flux, *virtual_salt, *relax_salt, *real_salt_flux;
    double is_nonlinfs;
    int wsplit;   // use_wsplit: W = w_e + w_i, w_i taken implicitly (adv_tra_vert_impl, implicit diffusion)
    const double *Ki;      // N2*L horizontal/isoneutral diffusivity (o_ARRAYS Ki)
    const double *slope;   // N2*L*3 slope_tapered, component fastest
    int hdiff;             // Ki set: diff_part_hor_redi runs
    int redi;              // Redi: slope terms, diff_ver_part_redi_expl, K33 of the implicit diffusion
    // surface fluxes of the tracer IDs 101-104 and 304 (fgpu_set_surface_forcing_tracers), null until set
    const double *prec_rain;   // N2, ID 101
    const double *wiso;        // N2*3, node fastest (wiso_flux_oce(:,1..3)), IDs 102-104
    const double *hosing;      // N2, ID 304
    const double *hosing3d;    // N2*L, ID 304 in depth mode
    int hosing_depth;          // use_hosing .and. hosing_mode == 'depth'
};

// bc_surface (oce_ale_tracer.F90:1828-2108) for the IDs of a plain build; the
// ID and the presence of its flux array are checked on the host before any
// launch (fgpu_check_bc_inputs). nzmin is the column's first wet level.
__device__ __forceinline__ double bc_surface_dev(int id, double dt, const DevDyn &d, int N2, int L, int n, int nzmin,
                                                 double sval)
{
    switch (id) {
    case 1: return -dt * (d.heat_flux[n] / FGPU_VCPW + sval * d.water_flux[n] * d.is_nonlinfs);
    case 2: return dt * (d.virtual_salt[n] + d.relax_salt[n] + d.real_salt_flux[n] * d.is_nonlinfs);
    case 101: return dt * d.prec_rain[n];                                  // :2059-2060
    case 102: case 103: case 104: return dt * d.wiso[(size_t)(id - 102) * N2 + n];   // :2062-2067
    case 304:                                                               // :2074-2079
        return d.hosing_depth ? dt * d.hosing3d[(size_t)n * L + nzmin - 1] : dt * d.hosing[n];
    // ice-shelf water from basal melting, under a cavity only (:2080-2085)
    case 501: return nzmin == 1 ? 0.0 : dt * d.water_flux[n] * (sval - 1.0);
    default: return 0.0;   // 100, 301-303
    }
}
// the hosing tracer in depth mode takes bc_surface at every wet layer of the
// column, with that layer as the level being filled (oce_ale_tracer.F90:1374-1380);
// every other tracer at the first wet layer only
__device__ __forceinline__ bool bc_every_layer(int id, const DevDyn &d) { return id == 304 && d.hosing_depth; }

// Per-tracer state, tracer-major: tracer t occupies [t*stride, (t+1)*stride).
struct DevTracers {
    int ntr;
    const int *id, *abo, *hor, *ver, *lim, *ivd;   // ntr each
    const double *ph, *pv;                         // ntr each
    size_t sN, sNI, sE, sF2, sOld, sG;             // strides: N2*L, N2*NL, E2*L, F2*L*2, N2*L*(AB_order-1), N2*L*2
    size_t sStg;              // tracer stride of tr_xy/aux: F2*L*2 with a slot per tracer, 0 with one shared slot
    double *T, *Told;
    double *lo_h, *a_h;       // E2*L
    double *fct_lo;           // N2*L
    double *a_v;              // N2*NL
    double *rp, *rm;          // N2*L
    double *gmean;            // N2*L*2 node-mean gradients of the fused step
    // buffers of the per-stage kernels only, allocated on first use (fgpu_ensure_staged)
    double *TAB;              // N2*L per tracer (valuesAB)
    double *tr_xy, *aux;      // F2*L*2 per slot (tracers%work: one slot shared by the per-tracer calls)
    double *dttf_h, *dttf_v;  // N2*L, one tracer's worth
    double *dbg_av;           // N2*NL limited vertical flux of the last per-tracer call
    double *dbg_grad;         // E2*L*4 materialised edge_up_dn_grad of the last per-tracer call
    // explicit diffusion increments of del_ttf from T^n, per tracer (fgpu_diff.hip), allocated on use
    double *hd;               // E2*L: edge flux c of diff_part_hor_redi (node 1 adds 0+c, node 2 adds 0-c)
    double *ve;               // N2*L: diff_ver_part_expl_ale increment (i_vert_diff off)
    double *vr;               // N2*L: diff_ver_part_redi_expl increment
};

struct ProfEntry {
    std::string name;
    double total_ms = 0.0;
    int launches = 0;
};

// variants of the fused step's last kernel (same results bit for bit)
#define KD_LOCKSTEP 0    // kD_node: the block's waves in lock-step, sweeps on the first wave's lanes
#define KD_SWEEPWAVE 1   // kDp_node<..., true>: one wave per column plus a sweep wave per block
#define KD_WAVELOCAL 2   // kDp_node<..., false>: one wave per column, which sweeps its own column
// kDp_node's LDS rows for columns of 65..KDP_SHORT_NL interfaces (70-layer
// meshes): KDP_SHORT_NL + 2 doubles instead of 130, so that more blocks fit per
// CU, and 3 or 4 columns per block
#define KDP_SHORT_NL 72

struct fgpu_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool mesh_ready = false, vert_ready = false, vel_ready = false, kv_ready = false, forcing_ready = false;
    bool tracers_ready = false;
    DevMesh m{};
    DevDyn d{};
    DevTracers t{};
    int N2 = 0, F2 = 0, E2 = 0, nl = 0, kmax = 0;
    double cyclic_length = 0.0;
    std::vector<void *> allocs;        // everything hipMalloc'ed by the context
    std::vector<void *> tracer_allocs; // released on fgpu_tracers_init
    // tracer configuration (host mirror)
    std::vector<int> h_id, h_abo, h_hor, h_ver, h_lim, h_ivd;
    std::vector<double> h_ph, h_pv;
    std::vector<char> h_configured;
    int *d_kneed = nullptr;   // first level of a node whose mean gradient an edge may use
    int *d_cfg_i = nullptr;
    double *d_cfg_d = nullptr;
    // host copies needed for setup
    std::vector<int> h_updn, h_edges, h_edge_tri, h_ne_ptr, h_ne_idx, h_elem_nodes, h_ulevn;
    // step graph
    hipGraphExec_t graph = nullptr;
    double graph_dt = 0.0;
    bool graph_valid = false;
    bool fused = true;        // batched step: fused kernels (false: one kernel per reference stage)
    int step_tpb = 4096;      // tracers per block of the fused step kernels (clamped to the tracer count)
    int step_tch = 2;         // tracers per Thomas chunk of the fused step's last kernel (2 or 4)
    int step_cpb = 1;         // columns per block of the fused step's last kernel (1, 4 or 8; 64-lane columns)
    int step_order = 1;       // processing order of the fused step's columns: 1 Hilbert, 0 index (FGPU_STEP_ORDER)
    int step_kcf = 0;         // K_C over (tracer, level) items with this many tracers per block (1, 2, 4); 0: kC_node
    int step_ka2 = 0;         // 1: K_A with two adjacent levels per lane (kA2_edge), for an even number of layers
    int step_kb2 = 0;         // 1: K_B likewise (kB2_node), without the W split and CDIFF/PPM tracers
    int step_kdp = 0;         // last kernel of the plain step: KD_LOCKSTEP (kD_node), KD_SWEEPWAVE or KD_WAVELOCAL (kDp_node)
    int stage_slots = 0;      // tr_xy/aux slots allocated for the per-stage kernels (0: none yet)
    // profiling
    bool prof = false;
    std::vector<ProfEntry> prof_entries;
    struct Pending { int idx; hipEvent_t e0, e1; };
    std::vector<Pending> prof_pending;
    std::vector<hipEvent_t> ev_free;
    // reductions
    double *d_red = nullptr;   // scratch for deterministic reductions
    // SSH
    int ssh_n = 0, ssh_nnz = 0, ssh_variant = -1;
    bool ssh_ready = false, ssh_pr_ready = false;
    int *ssh_rowptr = nullptr, *ssh_colind = nullptr;
    double *ssh_val = nullptr, *ssh_pr = nullptr, *ssh_diag_inv = nullptr;
    double *ssh_x = nullptr, *ssh_x0 = nullptr, *ssh_rhs = nullptr;
    double *ssh_r = nullptr, *ssh_z = nullptr, *ssh_p = nullptr, *ssh_p1 = nullptr, *ssh_q = nullptr;
    double *ssh_rec0 = nullptr, *ssh_rec1 = nullptr;   // single-reduction CG: 4 doubles per row {r, w, s, D^-1}, per parity
    bool ssh_single_reduction = true;   // diagonal preconditioner: one reduction (one kernel) per iteration
    hipGraphExec_t ssh_graph = nullptr;
    int ssh_graph_variant = -1, ssh_last_iters = 0;
    double *ssh_part = nullptr;   // per-block partial sums
    double *ssh_out = nullptr;    // device stats
    int ssh_blocks = 0;
    std::vector<int> h_ssh_rowptr, h_ssh_colind;
    // learned preconditioner (fgpu_mpnn.hip)
    struct NNDev *nn = nullptr;
    // vert_vel_ale / compute_ssh_rhs_ale on the device (fgpu_dyn.hip)
    double *vv_hbar = nullptr, *vv_hbar_old = nullptr, *vv_cfl = nullptr;
    double *vv_uvr = nullptr, *vv_old = nullptr, *vv_ce = nullptr, *vv_rhs = nullptr;
    double *vv_dhe = nullptr;       // dhe of update_stiff_mat_ale (elem2D)
    bool hbar_valid = false;        // vv_hbar, vv_hbar_old, vv_old (ssh_rhs_old), vv_dhe set by compute_hbar_ale
    double *ale_zbar = nullptr;     // zlevel: mesh%zbar (nl), min_hnode and lzstar_lev (fgpu_set_ale_config)
    double ale_min_hnode = 0.5;
    int ale_lzstar = 4;
    double *vv_colmax = nullptr;    // per-column max of CFL_z (nod2D)
    void *vv_flags = nullptr;       // max CFL_z bits, NaN column count
};

// error helpers (fgpu_runtime.hip)
int fgpu_fail(int code, const std::string &msg);
#define FGPU_HIP(call)                                                               \
    do {                                                                             \
        hipError_t _e = (call);                                                      \
        if (_e != hipSuccess)                                                        \
            return fgpu_fail(FGPU_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

int fgpu_alloc(fgpu_context *c, void **p, size_t bytes, bool tracer_scope = false);
// FGPU_ERR_STATE unless every flux array that the IDs of tracers tr0..tr0+ntr-1 read is set
int fgpu_check_bc_inputs(fgpu_context *c, int tr0, int ntr);

// launch bracketing for the per-kernel profile (fgpu_runtime.hip): when the
// context profiles, the constructor records an event before the launch and the
// destructor one after it; fgpu_prof_flush() adds the elapsed times up.
struct ProfScope {
    fgpu_context *c;
    int idx = -1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProfScope(fgpu_context *c, const char *name);
    ~ProfScope();
};
int fgpu_prof_flush(fgpu_context *c);

// fused batched step (fgpu_step.hip): four launches per step for all tracers
int fgpu_enqueue_fused_step(fgpu_context *c, double dt);
void fgpu_kd_shape(const fgpu_context *c, int *tch, int *cpb);
int fgpu_kd_extras(const fgpu_context *c);
int fgpu_compute_zn(fgpu_context *c);
int fgpu_ensure_staged(fgpu_context *c, int slots);
// explicit diffusion terms (fgpu_diff.hip)
int fgpu_ensure_diff(fgpu_context *c);
int fgpu_launch_diff(fgpu_context *c, int tr0, int ntr, double dt);

// SSH solver pieces shared with the learned preconditioner
struct FgpuCgSummary {
    int iters, done, nbreak, nneg;   // done: 1 converged, 2 breakdown, 3 iteration cap
    double rr_last, rtol;
};
int fgpu_ssh_cg_from_x(fgpu_context *c, double soltol, int maxiter, int variant, FgpuCgSummary *out);
void fgpu_nn_release(fgpu_context *c);
