// Context, mesh upload and data movement of libfesom_gpu.so (gfx950).
//
// Mesh preprocessing done once at upload, as the reference does in mesh_setup
// and muscl_adv_init:
//   * 1-based -> 0-based connectivity;
//   * per-edge level numbers nl1/nu1/nl2/nu2 (oce_adv_tra_hor.F90:104-128) and
//     the nu12/nl12 max/min rule of the edge scatters (oce_adv_tra_driver.F90:142-156);
//   * node -> edge gather lists in ascending edge order, so that every
//     edge->node sum is a deterministic gather in the serial reference order
//     (no float atomics);
//   * up/downwind triangles (oce_muscl_adv.F90:162-352);
//   * SSH sparsity with the diagonal first (oce_ale.F90:1440-1496).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include <algorithm>
#include <cstdint>

#include "fgpu_internal.h"

static thread_local std::string g_last_error;

int fgpu_fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

extern "C" const char *fgpu_last_error(void) { return g_last_error.c_str(); }
extern "C" int fgpu_version(void) { return 1; }

int fgpu_alloc(fgpu_context *c, void **p, size_t bytes, bool tracer_scope)
{
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess)
        return fgpu_fail(FGPU_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                             "): " + hipGetErrorString(e));
    (tracer_scope ? c->tracer_allocs : c->allocs).push_back(*p);
    return FGPU_OK;
}

template <typename T>
static int upload(fgpu_context *c, T **dst, const T *src, size_t n)
{
    int rc = fgpu_alloc(c, (void **)dst, n * sizeof(T));
    if (rc) return rc;
    FGPU_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return FGPU_OK;
}

// ---------------------------------------------------------------- profiling
ProfScope::ProfScope(fgpu_context *cc, const char *name) : c(cc)
{
    if (!c->prof) return;
    for (size_t i = 0; i < c->prof_entries.size(); ++i)
        if (c->prof_entries[i].name == name) { idx = (int)i; break; }
    if (idx < 0) {
        ProfEntry pe;
        pe.name = name;
        c->prof_entries.push_back(pe);
        idx = (int)c->prof_entries.size() - 1;
    }
    for (hipEvent_t *e : {&e0, &e1}) {
        if (!c->ev_free.empty()) { *e = c->ev_free.back(); c->ev_free.pop_back(); }
        else if (hipEventCreate(e) != hipSuccess) { *e = nullptr; }
    }
    if (e0) (void)hipEventRecord(e0, c->stream);
}

ProfScope::~ProfScope()
{
    if (!c->prof || idx < 0 || !e0 || !e1) return;
    (void)hipEventRecord(e1, c->stream);
    c->prof_pending.push_back({idx, e0, e1});
}

int fgpu_prof_flush(fgpu_context *c)
{
    if (c->prof_pending.empty()) return FGPU_OK;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    for (auto &p : c->prof_pending) {
        float ms = 0.f;
        FGPU_HIP(hipEventElapsedTime(&ms, p.e0, p.e1));
        c->prof_entries[p.idx].total_ms += ms;
        c->prof_entries[p.idx].launches += 1;
        c->ev_free.push_back(p.e0);
        c->ev_free.push_back(p.e1);
    }
    c->prof_pending.clear();
    return FGPU_OK;
}

extern "C" int fgpu_profile_enable(fgpu_context *c, int on)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    int rc = fgpu_prof_flush(c);
    if (rc) return rc;
    c->prof = on != 0;
    c->prof_entries.clear();
    return FGPU_OK;
}

extern "C" int fgpu_profile_count(fgpu_context *c)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    int rc = fgpu_prof_flush(c);
    if (rc) return rc;
    return (int)c->prof_entries.size();
}

extern "C" int fgpu_profile_read(fgpu_context *c, int idx, char *name, int name_len, double *total_ms,
                                 int *launches)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    int rc = fgpu_prof_flush(c);
    if (rc) return rc;
    if (idx < 0 || idx >= (int)c->prof_entries.size()) return fgpu_fail(FGPU_ERR_ARG, "profile index");
    const ProfEntry &e = c->prof_entries[idx];
    if (name && name_len > 0) {
        std::strncpy(name, e.name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (total_ms) *total_ms = e.total_ms;
    if (launches) *launches = e.launches;
    return FGPU_OK;
}

// ---------------------------------------------------------------- context
extern "C" int fgpu_create(fgpu_context **out, int device)
{
    if (!out) return fgpu_fail(FGPU_ERR_ARG, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    FGPU_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fgpu_fail(FGPU_ERR_ARG, "device " + std::to_string(device) + " out of range (" +
                                           std::to_string(ndev) + " visible)");
    FGPU_HIP(hipSetDevice(device));
    fgpu_context *c = new fgpu_context();
    c->device = device;
    const char *sv = std::getenv("FGPU_STEP");
    c->fused = !(sv && std::string(sv) == "staged");
    if (const char *bv = std::getenv("FGPU_STEP_TPB")) {
        int v = std::atoi(bv);
        if (v >= 1 && v <= 4096) c->step_tpb = v;
    }
    if (const char *cv = std::getenv("FGPU_SSH_CG")) c->ssh_single_reduction = std::string(cv) != "classic";
    if (const char *tv = std::getenv("FGPU_STEP_TCH")) c->step_tch = std::atoi(tv);
    if (const char *pv = std::getenv("FGPU_STEP_CPB")) c->step_cpb = std::atoi(pv);
    if (const char *ov = std::getenv("FGPU_STEP_ORDER")) c->step_order = std::string(ov) != "index";
    if (const char *fv = std::getenv("FGPU_STEP_KC")) c->step_kcf = std::atoi(fv);
    if (const char *av = std::getenv("FGPU_STEP_KA")) c->step_ka2 = std::atoi(av) == 2;
    if (const char *bv = std::getenv("FGPU_STEP_KB")) c->step_kb2 = std::atoi(bv) == 2;
    if (const char *kv = std::getenv("FGPU_STEP_KD")) {
        const std::string v(kv);
        c->step_kdp = v == "sweepwave" ? KD_SWEEPWAVE : (v == "wavelocal" ? KD_WAVELOCAL : KD_LOCKSTEP);
    }
    // default priority: stream priorities are left alone on purpose
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fgpu_fail(FGPU_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return FGPU_OK;
}

static void free_list(std::vector<void *> &v)
{
    for (void *p : v) (void)hipFree(p);
    v.clear();
}

// free one tracer-scope allocation (the stream must be idle)
static void free_tracer_buf(fgpu_context *c, double *&p)
{
    if (!p) return;
    auto &v = c->tracer_allocs;
    for (size_t i = 0; i < v.size(); ++i)
        if (v[i] == (void *)p) { v.erase(v.begin() + (long)i); break; }
    (void)hipFree(p);
    p = nullptr;
}

// Per-stage buffers: valuesAB for every tracer, the del_ttf/limited-flux/edge
// gradient work arrays of one tracer, and `slots` copies of tr_xy and the FCT
// element bounds (1 for the per-tracer entry points, which share them like
// tracers%work; ntr for the batched per-stage step). Allocated on first use so
// that the fused step's footprint stays at what it reads and writes.
int fgpu_ensure_staged(fgpu_context *c, int slots)
{
    DevTracers &t = c->t;
    if (t.TAB && c->stage_slots >= slots) return FGPU_OK;
    FGPU_HIP(hipSetDevice(c->device));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    int rc;
    if (!t.TAB) {
        struct A { double **p; size_t n; } arrs[] = {
            {&t.TAB, t.sN * (size_t)t.ntr}, {&t.dttf_h, t.sN}, {&t.dttf_v, t.sN}, {&t.dbg_av, t.sNI},
            {&t.dbg_grad, t.sE * 4}};
        for (auto &a : arrs) {
            if ((rc = fgpu_alloc(c, (void **)a.p, a.n * sizeof(double), true))) return rc;
            FGPU_HIP(hipMemset(*a.p, 0, a.n * sizeof(double)));
        }
    }
    if (c->stage_slots < slots) {
        free_tracer_buf(c, t.tr_xy);
        free_tracer_buf(c, t.aux);
        for (double **p : {&t.tr_xy, &t.aux}) {
            if ((rc = fgpu_alloc(c, (void **)p, t.sF2 * (size_t)slots * sizeof(double), true))) return rc;
            FGPU_HIP(hipMemset(*p, 0, t.sF2 * (size_t)slots * sizeof(double)));
        }
        c->stage_slots = slots;
        t.sStg = slots > 1 ? t.sF2 : 0;
    }
    c->graph_valid = false;
    return FGPU_OK;
}

extern "C" int fgpu_destroy(fgpu_context *c)
{
    if (!c) return FGPU_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &p : c->prof_pending) { (void)hipEventDestroy(p.e0); (void)hipEventDestroy(p.e1); }
    for (auto e : c->ev_free) (void)hipEventDestroy(e);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->ssh_graph) (void)hipGraphExecDestroy(c->ssh_graph);
    fgpu_nn_release(c);
    free_list(c->tracer_allocs);
    free_list(c->allocs);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return FGPU_OK;
}

extern "C" int fgpu_synchronize(fgpu_context *c)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

// ---------------------------------------------------------------- mesh
static double trim_cyc(double b, double Lc)
{   // gen_modules_rotate_grid.F90:204-211
    if (b > Lc / 2.0) b = b - Lc;
    if (b < -Lc / 2.0) b = b + Lc;
    return b;
}

// find_up_downwind_triangles (oce_muscl_adv.F90:162-352), one search at a node;
// every matching branch is a CYCLE in the reference, so the last match wins.
static int updn_search(const double *coord, const int *elem_nodes, const int *nie, const int *nie_num,
                       int kmax, double Lc, int node, const double x[2])
{
    int found = -1;
    for (int k = 0; k < nie_num[node]; ++k) {
        int elem = nie[(size_t)node * kmax + k];
        const int *en = elem_nodes + 3 * (size_t)elem;
        int a0, a1, a2;
        if (en[0] == node) { a0 = 0; a1 = 1; a2 = 2; }
        else if (en[1] == node) { a0 = 1; a1 = 0; a2 = 2; }
        else { a0 = 2; a1 = 0; a2 = 1; }
        double b[2], c[2];
        b[0] = coord[2 * en[a1]] - coord[2 * en[a0]];
        b[1] = coord[2 * en[a1] + 1] - coord[2 * en[a0] + 1];
        c[0] = coord[2 * en[a2]] - coord[2 * en[a0]];
        c[1] = coord[2 * en[a2] + 1] - coord[2 * en[a0] + 1];
        b[0] = trim_cyc(b[0], Lc);
        c[0] = trim_cyc(c[0], Lc);
        double cr = c[0] * c[0] + c[1] * c[1];
        double bx = (b[0] * c[0] + b[1] * c[1]) / cr;
        double by = (-b[0] * c[1] + b[1] * c[0]) / cr;
        double xx = (x[0] * c[0] + x[1] * c[1]) / cr;
        double xy = (-x[0] * c[1] + x[1] * c[0]) / cr;
        double ab = std::atan2(by, bx);
        double ax = std::atan2(xy, xx);
        if ((ab > 0.0) && (ax > 0.0) && (ax < ab)) { found = elem; continue; }
        if ((ab < 0.0) && (ax < 0.0) && (ax > ab)) { found = elem; continue; }
        if ((ab == ax) || (ax == 0.0)) { found = elem; continue; }
    }
    return found;
}

// index along a Hilbert curve of the cell (x, y) of an n x n grid, n a power
// of two (the iterative form of Hilbert's quadrant recursion)
static uint64_t hilbert_index(uint32_t n, uint32_t x, uint32_t y)
{
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {   // rotate the quadrant so the curve's pieces join
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

// points ordered along a Hilbert curve over their bounding box (65536^2 cells)
static std::vector<int> hilbert_order(const std::vector<double> &x, const std::vector<double> &y)
{
    const size_t n = x.size();
    std::vector<int> ord(n);
    if (n == 0) return ord;
    const double x0 = *std::min_element(x.begin(), x.end()), x1 = *std::max_element(x.begin(), x.end());
    const double y0 = *std::min_element(y.begin(), y.end()), y1 = *std::max_element(y.begin(), y.end());
    const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
    std::vector<std::pair<uint64_t, int>> key(n);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t cx = (uint32_t)std::min(65535.0, std::max(0.0, (x[i] - x0) * sx));
        const uint32_t cy = (uint32_t)std::min(65535.0, std::max(0.0, (y[i] - y0) * sy));
        key[i] = {hilbert_index(65536u, cx, cy), (int)i};
    }
    std::sort(key.begin(), key.end());
    for (size_t i = 0; i < n; ++i) ord[i] = key[i].second;
    return ord;
}

static std::vector<int> identity_order(int n)
{
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    return ord;
}

extern "C" int fgpu_mesh_init(fgpu_context *c, int nod2D, int elem2D, int edge2D, int nl, int kmax,
                              double cyclic_length, const double *coord_nod2D, const int *elem2D_nodes,
                              const int *edges, const int *edge_tri, const int *nod_in_elem2D,
                              const int *nod_in_elem2D_num, const int *nlevels, const int *ulevels,
                              const int *nlevels_nod2D, const int *ulevels_nod2D,
                              const int *nlevels_nod2D_min, const int *ulevels_nod2D_max,
                              const double *elem_area, const double *elem_cos, const double *edge_dxdy,
                              const double *edge_cross_dxdy, const double *gradient_sca)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh already initialised for this context");
    if (nod2D <= 0 || elem2D <= 0 || edge2D <= 0 || kmax <= 0)
        return fgpu_fail(FGPU_ERR_ARG, "empty mesh");
    if (nl < 3 || nl > FGPU_MAX_NL) return fgpu_fail(FGPU_ERR_ARG, "nl out of range [3, 128]");
    if (!coord_nod2D || !elem2D_nodes || !edges || !edge_tri || !nod_in_elem2D || !nod_in_elem2D_num ||
        !nlevels || !ulevels || !nlevels_nod2D || !ulevels_nod2D || !nlevels_nod2D_min ||
        !ulevels_nod2D_max || !elem_area || !elem_cos || !edge_dxdy || !edge_cross_dxdy || !gradient_sca)
        return fgpu_fail(FGPU_ERR_ARG, "null mesh array");
    FGPU_HIP(hipSetDevice(c->device));
    const int N2 = nod2D, F2 = elem2D, E2 = edge2D, L = nl - 1;
    // ---- validate and convert connectivity to 0-based
    std::vector<int> en(3 * (size_t)F2), ed(2 * (size_t)E2), et(2 * (size_t)E2), nie((size_t)N2 * kmax);
    for (size_t i = 0; i < en.size(); ++i) {
        en[i] = elem2D_nodes[i] - 1;
        if (en[i] < 0 || en[i] >= N2) return fgpu_fail(FGPU_ERR_ARG, "elem2D_nodes out of range");
    }
    for (size_t i = 0; i < ed.size(); ++i) {
        ed[i] = edges[i] - 1;
        if (ed[i] < 0 || ed[i] >= N2) return fgpu_fail(FGPU_ERR_ARG, "edges out of range");
    }
    for (int e = 0; e < E2; ++e) {
        int a = edge_tri[2 * e], b = edge_tri[2 * e + 1];
        if (a < 1 || a > F2) return fgpu_fail(FGPU_ERR_ARG, "edge_tri(1,:) out of range");
        if (b > F2) return fgpu_fail(FGPU_ERR_ARG, "edge_tri(2,:) out of range");
        et[2 * e] = a - 1;
        et[2 * e + 1] = b > 0 ? b - 1 : -1;
    }
    for (int n = 0; n < N2; ++n) {
        int num = nod_in_elem2D_num[n];
        if (num < 1 || num > kmax) return fgpu_fail(FGPU_ERR_ARG, "nod_in_elem2D_num out of range");
        for (int k = 0; k < kmax; ++k) {
            int v = nod_in_elem2D[(size_t)n * kmax + k];
            if (k < num) {
                if (v < 1 || v > F2) return fgpu_fail(FGPU_ERR_ARG, "nod_in_elem2D out of range");
                nie[(size_t)n * kmax + k] = v - 1;
            } else {
                nie[(size_t)n * kmax + k] = -1;
            }
        }
    }
    for (int e = 0; e < F2; ++e)
        if (ulevels[e] < 1 || nlevels[e] - ulevels[e] < 1 || nlevels[e] > nl)
            return fgpu_fail(FGPU_ERR_ARG, "element level numbers out of range");
    for (int n = 0; n < N2; ++n)
        if (ulevels_nod2D[n] < 1 || nlevels_nod2D[n] - ulevels_nod2D[n] < 2 || nlevels_nod2D[n] > nl ||
            nlevels_nod2D_min[n] > nlevels_nod2D[n] || nlevels_nod2D_min[n] < ulevels_nod2D[n] + 1 ||
            ulevels_nod2D_max[n] < ulevels_nod2D[n])
            return fgpu_fail(FGPU_ERR_ARG, "node level numbers out of range (columns need >= 2 layers)");
    // ---- per-edge level numbers
    std::vector<int> einfo(4 * (size_t)E2), erange(2 * (size_t)E2);
    for (int e = 0; e < E2; ++e) {
        int el1 = et[2 * e], el2 = et[2 * e + 1];
        int nl1 = nlevels[el1] - 1, nu1 = ulevels[el1], nl2 = 0, nu2 = 0;
        if (el2 >= 0) { nl2 = nlevels[el2] - 1; nu2 = ulevels[el2]; }
        einfo[4 * e + 0] = nl1; einfo[4 * e + 1] = nu1; einfo[4 * e + 2] = nl2; einfo[4 * e + 3] = nu2;
        int nl12 = std::max(nl1, nl2), nu12 = nu1;
        if (nu2 > 0) nu12 = std::min(nu1, nu2);
        erange[2 * e] = nu12;
        erange[2 * e + 1] = nl12;
        int n1 = ed[2 * e], n2 = ed[2 * e + 1];
        if (nl12 > nlevels_nod2D[n1] - 1 || nl12 > nlevels_nod2D[n2] - 1)
            return fgpu_fail(FGPU_ERR_ARG, "edge deeper than one of its nodes");
    }
    // ---- node -> edge gather lists, edges ascending
    std::vector<int> ne_ptr(N2 + 1, 0), ne_idx(2 * (size_t)E2);
    for (int e = 0; e < E2; ++e) { ne_ptr[ed[2 * e] + 1]++; ne_ptr[ed[2 * e + 1] + 1]++; }
    for (int n = 0; n < N2; ++n) ne_ptr[n + 1] += ne_ptr[n];
    {
        std::vector<int> fill(ne_ptr.begin(), ne_ptr.end() - 1);
        for (int e = 0; e < E2; ++e) {
            ne_idx[fill[ed[2 * e]]++] = 2 * e;
            ne_idx[fill[ed[2 * e + 1]]++] = 2 * e + 1;
        }
    }
    {
        int maxdeg = 0;
        for (int n = 0; n < N2; ++n) maxdeg = std::max(maxdeg, ne_ptr[n + 1] - ne_ptr[n]);
        if (maxdeg > 64 || kmax > 64) return fgpu_fail(FGPU_ERR_ARG, "node degree above 64");
    }
    // ---- up/downwind triangles
    std::vector<int> updn(2 * (size_t)E2);
    for (int e = 0; e < E2; ++e) {
        int n1 = ed[2 * e], n2 = ed[2 * e + 1];
        double x[2];
        x[0] = coord_nod2D[2 * n2] - coord_nod2D[2 * n1];
        x[1] = coord_nod2D[2 * n2 + 1] - coord_nod2D[2 * n1 + 1];
        if (x[0] > cyclic_length / 2.0) x[0] = x[0] - cyclic_length;
        if (x[0] < -cyclic_length / 2.0) x[0] = x[0] + cyclic_length;
        double xm[2] = {-x[0], -x[1]};
        updn[2 * e] = updn_search(coord_nod2D, en.data(), nie.data(), nod_in_elem2D_num, kmax,
                                  cyclic_length, n1, xm);
        updn[2 * e + 1] = updn_search(coord_nod2D, en.data(), nie.data(), nod_in_elem2D_num, kmax,
                                      cyclic_length, n2, x);
    }
    // ---- lowest level at which an edge needs the mean gradient of a node
    //      (the non-shared levels of fill_up_dn_grad, oce_muscl_adv.F90:383-520)
    std::vector<int> kneed(N2);
    for (int n = 0; n < N2; ++n) kneed[n] = nlevels_nod2D[n];
    for (int e = 0; e < E2; ++e) {
        int n1 = ed[2 * e], n2 = ed[2 * e + 1];
        bool both = updn[2 * e] >= 0 && updn[2 * e + 1] >= 0;
        int nzmin = std::max(ulevels_nod2D_max[n1], ulevels_nod2D_max[n2]);
        int nzmax = std::min(nlevels_nod2D_min[n1], nlevels_nod2D_min[n2]);
        for (int nn : {n1, n2}) {
            int lo;
            if (!both || ulevels_nod2D[nn] <= nzmin - 1) lo = ulevels_nod2D[nn];
            else lo = nzmax;
            kneed[nn] = std::min(kneed[nn], lo);
        }
    }
    // ---- upload
    DevMesh &m = c->m;
    m.N2 = N2; m.F2 = F2; m.E2 = E2; m.L = L; m.NL = nl; m.kmax = kmax;
    int rc;
#define UP(dst, src, n)                                                \
    do {                                                               \
        rc = upload(c, &dst, src, (size_t)(n));                        \
        if (rc) return rc;                                             \
    } while (0)
    int *p_i;
    double *p_d;
    UP(p_i, en.data(), en.size()); m.elem_nodes = p_i;
    UP(p_i, ed.data(), ed.size()); m.edges = p_i;
    UP(p_i, et.data(), et.size()); m.edge_tri = p_i;
    UP(p_i, einfo.data(), einfo.size()); m.einfo = p_i;
    UP(p_i, erange.data(), erange.size()); m.erange = p_i;
    UP(p_i, updn.data(), updn.size()); m.updn = p_i;
    UP(p_i, nie.data(), nie.size()); m.nie = p_i;
    UP(p_i, nod_in_elem2D_num, N2); m.nie_num = p_i;
    UP(p_i, nlevels, F2); m.nlev = p_i;
    UP(p_i, ulevels, F2); m.ulev = p_i;
    UP(p_i, nlevels_nod2D, N2); m.nlevn = p_i;
    UP(p_i, ulevels_nod2D, N2); m.ulevn = p_i;
    UP(p_i, nlevels_nod2D_min, N2); m.nlevn_min = p_i;
    UP(p_i, ulevels_nod2D_max, N2); m.ulevn_max = p_i;
    UP(p_i, ne_ptr.data(), ne_ptr.size()); m.ne_ptr = p_i;
    UP(p_i, ne_idx.data(), ne_idx.size()); m.ne_idx = p_i;
    {
        // per-node gather records: the node's edge list with each edge's level
        // ranges, so a node kernel reaches them with one dependent load
        std::vector<int> rec(8 * ne_idx.size()), nrec(4 * (size_t)N2);
        for (size_t p = 0; p < ne_idx.size(); ++p) {
            const int q = ne_idx[p], e = q >> 1;
            int *r = rec.data() + 8 * p;
            r[0] = q;
            r[1] = ed[2 * e + (1 - (q & 1))];
            r[2] = erange[2 * e];
            r[3] = erange[2 * e + 1];
            for (int j = 0; j < 4; ++j) r[4 + j] = einfo[4 * e + j];
        }
        for (int n = 0; n < N2; ++n) {
            nrec[4 * n] = ne_ptr[n];
            nrec[4 * n + 1] = ne_ptr[n + 1] - ne_ptr[n];
            nrec[4 * n + 2] = ulevels_nod2D[n];
            nrec[4 * n + 3] = nlevels_nod2D[n];
        }
        UP(p_i, rec.data(), rec.size()); m.ne_rec = p_i;
        UP(p_i, nrec.data(), nrec.size()); m.node_rec = p_i;
    }
    UP(p_i, kneed.data(), kneed.size()); c->d_kneed = p_i;
    {
        // processing order of the fused step's column blocks: nodes by the
        // Hilbert index of their coordinates, edges by that of their midpoints
        std::vector<double> ex(E2), ey(E2);
        for (int e = 0; e < E2; ++e) {
            const int n1 = ed[2 * e], n2 = ed[2 * e + 1];
            ex[e] = 0.5 * (coord_nod2D[2 * n1] + coord_nod2D[2 * n2]);
            ey[e] = 0.5 * (coord_nod2D[2 * n1 + 1] + coord_nod2D[2 * n2 + 1]);
        }
        std::vector<double> nx(N2), ny(N2);
        for (int n = 0; n < N2; ++n) {
            nx[n] = coord_nod2D[2 * n];
            ny[n] = coord_nod2D[2 * n + 1];
        }
        const bool hil = c->step_order != 0;
        const std::vector<int> no = hil ? hilbert_order(nx, ny) : identity_order(N2);
        const std::vector<int> eo = hil ? hilbert_order(ex, ey) : identity_order(E2);
        UP(p_i, no.data(), no.size()); m.node_ord = p_i;
        UP(p_i, eo.data(), eo.size()); m.edge_ord = p_i;
    }
    // nboundary_lay of MUSCL (muscl_adv_init, oce_muscl_adv.F90:91-157): nl-1,
    // reset to 0 at both nodes of a boundary edge, else the minimum over the
    // node's edges of min(nlevels of the two triangles) - 1
    std::vector<int> nblay(N2, nl - 1);
    for (int e = 0; e < E2; ++e) {
        const int n1 = ed[2 * e], n2 = ed[2 * e + 1], el1 = et[2 * e], el2 = et[2 * e + 1];
        if (el1 < 0 || el2 < 0) {
            nblay[n1] = 0;
            nblay[n2] = 0;
        } else {
            const int v = std::min(nlevels[el1], nlevels[el2]) - 1;
            nblay[n1] = std::min(nblay[n1], v);
            nblay[n2] = std::min(nblay[n2], v);
        }
    }
    UP(p_i, nblay.data(), nblay.size()); m.nblay = p_i;
    {
        // FCT padding range per node from its elements (oce_adv_tra_fct.F90:148-178,
        // 195-215): an element without level k enters the node's bounds as
        // -1e3 / +1e3; some element around n lacks level k iff k < max ulevels
        // or k >= min nlevels over the elements around n
        std::vector<int> pad(2 * (size_t)N2);
        for (int n = 0; n < N2; ++n) {
            int umax = 1, nmin = nl + 1;
            for (int k = 0; k < nod_in_elem2D_num[n]; ++k) {
                const int el = nie[(size_t)n * kmax + k];
                umax = std::max(umax, ulevels[el]);
                nmin = std::min(nmin, nlevels[el]);
            }
            pad[2 * n] = umax;
            pad[2 * n + 1] = nmin;
        }
        UP(p_i, pad.data(), pad.size()); m.fct_pad = p_i;
    }
    {
        // packed edge records (mesh arrays and the fill_up_dn_grad level rule)
        std::vector<int> ri((size_t)E2 * ERI, 0);
        std::vector<double> rd((size_t)E2 * ERD, 0.0);
        for (int e = 0; e < E2; ++e) {
            int *r = ri.data() + (size_t)e * ERI;
            double *q = rd.data() + (size_t)e * ERD;
            const int n1 = ed[2 * e], n2 = ed[2 * e + 1], el1 = et[2 * e], el2 = et[2 * e + 1];
            r[ER_N1] = n1; r[ER_N2] = n2; r[ER_EL1] = el1; r[ER_EL2] = el2;
            for (int j = 0; j < 4; ++j) r[ER_NL1 + j] = einfo[4 * e + j];
            const int up = updn[2 * e], dn = updn[2 * e + 1];
            r[ER_RULE] = (up >= 0 && dn >= 0) ? 1 : 0;
            if (r[ER_RULE]) {
                for (int j = 0; j < 3; ++j) {
                    r[ER_UA0 + j] = en[3 * up + j];
                    r[ER_DA0 + j] = en[3 * dn + j];
                    q[ER_GU0 + j] = gradient_sca[6 * up + j];
                    q[ER_GU3 + j] = gradient_sca[6 * up + 3 + j];
                    q[ER_GD0 + j] = gradient_sca[6 * dn + j];
                    q[ER_GD3 + j] = gradient_sca[6 * dn + 3 + j];
                }
                r[ER_NZMIN] = std::max(ulevels_nod2D_max[n1], ulevels_nod2D_max[n2]);
                r[ER_NZMAX] = std::min(nlevels_nod2D_min[n1], nlevels_nod2D_min[n2]);
            }
            r[ER_U1] = ulevels_nod2D[n1]; r[ER_U2] = ulevels_nod2D[n2];
            r[ER_NL1N] = nlevels_nod2D[n1]; r[ER_NL2N] = nlevels_nod2D[n2];
            r[ER_NBL1] = nblay[n1]; r[ER_NBL2] = nblay[n2];
            q[ER_DX1] = edge_cross_dxdy[4 * e]; q[ER_DY1] = edge_cross_dxdy[4 * e + 1];
            double a = 6367500.0 * elem_cos[el1];
            if (el2 >= 0) {
                q[ER_DX2] = edge_cross_dxdy[4 * e + 2];
                q[ER_DY2] = edge_cross_dxdy[4 * e + 3];
                a = 0.5 * (a + 6367500.0 * elem_cos[el2]);   // oce_adv_tra_hor.F90:598-612
            }
            q[ER_A] = a;
            q[ER_EDX] = edge_dxdy[2 * e]; q[ER_EDY] = edge_dxdy[2 * e + 1];
        }
        UP(p_i, ri.data(), ri.size()); m.erec_i = p_i;
        UP(p_d, rd.data(), rd.size()); m.erec_d = p_d;
    }
    UP(p_d, elem_area, F2); m.elem_area = p_d;
    UP(p_d, elem_cos, F2); m.elem_cos = p_d;
    UP(p_d, edge_dxdy, 2 * (size_t)E2); m.edge_dxdy = p_d;
    UP(p_d, edge_cross_dxdy, 4 * (size_t)E2); m.ecd = p_d;
    UP(p_d, gradient_sca, 6 * (size_t)F2); m.gsca = p_d;
#undef UP
    c->N2 = N2; c->F2 = F2; c->E2 = E2; c->nl = nl; c->kmax = kmax;
    c->cyclic_length = cyclic_length;
    c->h_updn = updn;
    c->h_edges = ed;
    c->h_edge_tri = et;
    c->h_ne_ptr = ne_ptr;
    c->h_ne_idx = ne_idx;
    c->h_elem_nodes = en;
    c->h_ulevn.assign(ulevels_nod2D, ulevels_nod2D + N2);
    rc = fgpu_alloc(c, (void **)&c->d_red, sizeof(double) * 4096);
    if (rc) return rc;
    c->mesh_ready = true;
    return FGPU_OK;
}

extern "C" int fgpu_edge_up_dn_tri(fgpu_context *c, int *out)
{
    if (!c || !out) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    for (size_t i = 0; i < c->h_updn.size(); ++i) out[i] = c->h_updn[i] + 1;
    return FGPU_OK;
}

template <typename T>
static int upload_or_refresh(fgpu_context *c, const T **dst, const T *src, size_t n)
{
    if (!src) return fgpu_fail(FGPU_ERR_ARG, "null array");
    FGPU_HIP(hipStreamSynchronize(c->stream));   // no launch of the context still reads the old contents
    if (*dst == nullptr) {
        T *p;
        int rc = fgpu_alloc(c, (void **)&p, n * sizeof(T));
        if (rc) return rc;
        *dst = p;
    }
    FGPU_HIP(hipMemcpy((void *)*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return FGPU_OK;
}

extern "C" int fgpu_set_vertical(fgpu_context *c, const double *area, const double *areasvol,
                                 const double *hnode, const double *hnode_new, const double *helem,
                                 const double *zbar_3d_n, const double *Z_3d_n, const double *zbar_n_bot,
                                 const double *zbar_e_bot, const double *zbar_e_srf)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    FGPU_HIP(hipSetDevice(c->device));
    const size_t N2 = c->N2, F2 = c->F2, L = c->nl - 1, NL = c->nl;
    int rc;
    if ((rc = upload_or_refresh(c, &c->m.area, area, N2 * NL))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.areasvol, areasvol, N2 * NL))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.hnode, hnode, N2 * L))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.hnode_new, hnode_new, N2 * L))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.helem, helem, F2 * L))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.zbar3d, zbar_3d_n, N2 * NL))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.Z3d, Z_3d_n, N2 * L))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.zbar_n_bot, zbar_n_bot, N2))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.zbar_e_bot, zbar_e_bot, F2))) return rc;
    if ((rc = upload_or_refresh(c, &c->m.zbar_e_srf, zbar_e_srf, F2))) return rc;
    if ((rc = fgpu_compute_zn(c))) return rc;
    c->vert_ready = true;
    c->graph_valid = false;
    return FGPU_OK;
}

extern "C" int fgpu_set_velocity(fgpu_context *c, const double *uv, const double *w, const double *w_e,
                                 const double *w_i, int use_wsplit)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    FGPU_HIP(hipSetDevice(c->device));
    const size_t N2 = c->N2, F2 = c->F2, L = c->nl - 1, NL = c->nl;
    int rc;
    if ((rc = upload_or_refresh(c, &c->d.uv, uv, F2 * L * 2))) return rc;
    if ((rc = upload_or_refresh(c, &c->d.w, w, N2 * NL))) return rc;
    if ((rc = upload_or_refresh(c, &c->d.w_e, w_e, N2 * NL))) return rc;
    if ((rc = upload_or_refresh(c, &c->d.w_i, w_i, N2 * NL))) return rc;
    c->vel_ready = true;
    if (c->d.wsplit != (use_wsplit ? 1 : 0)) c->graph_valid = false;   // the step's kernels depend on it
    c->d.wsplit = use_wsplit ? 1 : 0;
    return FGPU_OK;
}

extern "C" int fgpu_set_ki(fgpu_context *c, const double *Ki, const double *slope_tapered, int redi)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    FGPU_HIP(hipSetDevice(c->device));
    c->graph_valid = false;
    if (!Ki) {   // Ki = 0: diff_part_hor_redi adds nothing
        if (redi) return fgpu_fail(FGPU_ERR_ARG, "Redi needs Ki");
        c->d.hdiff = 0;
        c->d.redi = 0;
        return FGPU_OK;
    }
    if (redi && !slope_tapered) return fgpu_fail(FGPU_ERR_ARG, "Redi needs slope_tapered");
    const size_t n = (size_t)c->N2 * (c->nl - 1);
    int rc = upload_or_refresh(c, &c->d.Ki, Ki, n);
    if (rc) return rc;
    if (slope_tapered && (rc = upload_or_refresh(c, &c->d.slope, slope_tapered, 3 * n))) return rc;
    c->d.hdiff = 1;
    c->d.redi = redi ? 1 : 0;
    return FGPU_OK;
}

extern "C" int fgpu_set_kv(fgpu_context *c, const double *Kv)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    FGPU_HIP(hipSetDevice(c->device));
    int rc = upload_or_refresh(c, &c->d.Kv, Kv, (size_t)c->N2 * c->nl);
    if (rc) return rc;
    c->kv_ready = true;
    return FGPU_OK;
}

extern "C" int fgpu_set_surface_forcing(fgpu_context *c, const double *heat_flux, const double *water_flux,
                                        const double *virtual_salt, const double *relax_salt,
                                        const double *real_salt_flux, double is_nonlinfs)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    FGPU_HIP(hipSetDevice(c->device));
    const size_t N2 = c->N2;
    int rc;
    if ((rc = upload_or_refresh(c, &c->d.heat_flux, heat_flux, N2))) return rc;
    if ((rc = upload_or_refresh(c, &c->d.water_flux, water_flux, N2))) return rc;
    if ((rc = upload_or_refresh(c, &c->d.virtual_salt, virtual_salt, N2))) return rc;
    if ((rc = upload_or_refresh(c, &c->d.relax_salt, relax_salt, N2))) return rc;
    if ((rc = upload_or_refresh(c, &c->d.real_salt_flux, real_salt_flux, N2))) return rc;
    c->d.is_nonlinfs = is_nonlinfs;
    c->forcing_ready = true;
    c->graph_valid = false;
    return FGPU_OK;
}

extern "C" int fgpu_set_surface_forcing_tracers(fgpu_context *c, const double *prec_rain,
                                                const double *wiso_flux_oce, const double *hosing_flux,
                                                const double *hosing_flux3D, int hosing_depth)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    FGPU_HIP(hipSetDevice(c->device));
    const size_t N2 = c->N2, L = c->nl - 1;
    int rc;
    if (prec_rain && (rc = upload_or_refresh(c, &c->d.prec_rain, prec_rain, N2))) return rc;
    if (wiso_flux_oce && (rc = upload_or_refresh(c, &c->d.wiso, wiso_flux_oce, 3 * N2))) return rc;
    if (hosing_flux && (rc = upload_or_refresh(c, &c->d.hosing, hosing_flux, N2))) return rc;
    if (hosing_flux3D && (rc = upload_or_refresh(c, &c->d.hosing3d, hosing_flux3D, N2 * L))) return rc;
    c->d.hosing_depth = hosing_depth ? 1 : 0;
    c->graph_valid = false;
    return FGPU_OK;
}

int fgpu_check_bc_inputs(fgpu_context *c, int tr0, int ntr)
{
    for (int i = tr0; i < tr0 + ntr; ++i) {
        const int id = c->h_id[i];
        const char *miss = nullptr;
        if (id == 101 && !c->d.prec_rain) miss = "prec_rain";
        if (id >= 102 && id <= 104 && !c->d.wiso) miss = "wiso_flux_oce";
        if (id == 304 && c->d.hosing_depth && !c->d.hosing3d) miss = "hosing_flux3D";
        if (id == 304 && !c->d.hosing_depth && !c->d.hosing) miss = "hosing_flux";
        if (miss)
            return fgpu_fail(FGPU_ERR_STATE, "tracer " + std::to_string(i + 1) + " (ID " + std::to_string(id) +
                                                 ") needs " + miss + " (fgpu_set_surface_forcing_tracers)");
    }
    return FGPU_OK;
}

// ---------------------------------------------------------------- tracers
extern "C" int fgpu_tracers_init(fgpu_context *c, int num_tracers)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    if (num_tracers < 1 || num_tracers > 4096) return fgpu_fail(FGPU_ERR_ARG, "num_tracers out of range");
    FGPU_HIP(hipSetDevice(c->device));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    free_list(c->tracer_allocs);
    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    c->graph_valid = false;
    const int nt = num_tracers;
    const size_t N2 = c->N2, F2 = c->F2, E2 = c->E2, L = c->nl - 1, NL = c->nl;
    DevTracers &t = c->t;
    t.ntr = nt;
    t.sN = N2 * L;
    t.sNI = N2 * NL;
    t.sE = E2 * L;
    t.sF2 = F2 * L * 2;
    t.sOld = N2 * L;   // AB2; grows to 2*N2*L when a tracer is configured with AB_order 3
    t.sG = N2 * L * 2;
    t.sStg = 0;
    t.TAB = t.tr_xy = t.aux = t.dttf_h = t.dttf_v = t.dbg_av = t.dbg_grad = nullptr;
    t.hd = t.ve = t.vr = nullptr;
    c->stage_slots = 0;
    int rc;
    // the fused batched step needs only these; the per-stage buffers come with
    // the first per-tracer call (fgpu_ensure_staged)
    struct A { double **p; size_t n; } arrs[] = {
        {&t.T, t.sN * nt}, {&t.Told, t.sOld * nt}, {&t.lo_h, t.sE * nt}, {&t.a_h, t.sE * nt},
        {&t.fct_lo, t.sN * nt}, {&t.a_v, t.sNI * nt}, {&t.rp, t.sN * nt}, {&t.rm, t.sN * nt},
        {&t.gmean, t.sG * nt}};
    for (auto &a : arrs) {
        if ((rc = fgpu_alloc(c, (void **)a.p, a.n * sizeof(double), true))) return rc;
        FGPU_HIP(hipMemset(*a.p, 0, a.n * sizeof(double)));
    }
    if ((rc = fgpu_alloc(c, (void **)&c->d_cfg_i, sizeof(int) * 6 * nt, true))) return rc;
    if ((rc = fgpu_alloc(c, (void **)&c->d_cfg_d, sizeof(double) * 2 * nt, true))) return rc;
    t.id = c->d_cfg_i;
    t.abo = c->d_cfg_i + nt;
    t.hor = c->d_cfg_i + 2 * nt;
    t.ver = c->d_cfg_i + 3 * nt;
    t.lim = c->d_cfg_i + 4 * nt;
    t.ivd = c->d_cfg_i + 5 * nt;
    t.ph = c->d_cfg_d;
    t.pv = c->d_cfg_d + nt;
    c->h_id.assign(nt, 1); c->h_abo.assign(nt, 2); c->h_hor.assign(nt, HOR_MFCT);
    c->h_ver.assign(nt, VER_QR4C); c->h_lim.assign(nt, LIM_FCT); c->h_ivd.assign(nt, 1);
    c->h_ph.assign(nt, 0.0); c->h_pv.assign(nt, 1.0);
    c->h_configured.assign(nt, 0);
    c->tracers_ready = true;
    return FGPU_OK;
}

static int parse_scheme(const char *s, const char *const *names, int nnames, int *code)
{
    if (!s) return -1;
    std::string v(s);
    while (!v.empty() && (v.back() == ' ' || v.back() == '\0')) v.pop_back();
    size_t b = 0;
    while (b < v.size() && v[b] == ' ') ++b;
    v = v.substr(b);
    for (int i = 0; i < nnames; ++i)
        if (v == names[i]) { *code = i; return 0; }
    return -1;
}

static int push_config(fgpu_context *c)
{
    const int nt = c->t.ntr;
    std::vector<int> ci(6 * nt);
    std::vector<double> cd(2 * nt);
    for (int i = 0; i < nt; ++i) {
        ci[i] = c->h_id[i]; ci[nt + i] = c->h_abo[i]; ci[2 * nt + i] = c->h_hor[i];
        ci[3 * nt + i] = c->h_ver[i]; ci[4 * nt + i] = c->h_lim[i]; ci[5 * nt + i] = c->h_ivd[i];
        cd[i] = c->h_ph[i]; cd[nt + i] = c->h_pv[i];
    }
    FGPU_HIP(hipMemcpy(c->d_cfg_i, ci.data(), ci.size() * sizeof(int), hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(c->d_cfg_d, cd.data(), cd.size() * sizeof(double), hipMemcpyHostToDevice));
    return FGPU_OK;
}

extern "C" int fgpu_tracer_config(fgpu_context *c, int tr_num, int ID, int AB_order, const char *tra_adv_hor,
                                  const char *tra_adv_ver, const char *tra_adv_lim, double tra_adv_ph,
                                  double tra_adv_pv, int i_vert_diff)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->tracers_ready) return fgpu_fail(FGPU_ERR_STATE, "tracers not initialised");
    if (tr_num < 1 || tr_num > c->t.ntr) return fgpu_fail(FGPU_ERR_ARG, "tr_num out of range");
    static const char *hor_names[] = {"MFCT", "UPW1", "MUSCL"};
    static const char *ver_names[] = {"QR4C", "UPW1", "CDIFF", "PPM"};
    static const char *lim_names[] = {"FCT", "NON"};
    int h, v, l;
    if (parse_scheme(tra_adv_hor, hor_names, 3, &h))
        return fgpu_fail(FGPU_ERR_SCHEME, std::string("Unknown horizontal advection type ") +
                                              (tra_adv_hor ? tra_adv_hor : "(null)") + "! Check your namelists!");
    if (parse_scheme(tra_adv_ver, ver_names, 4, &v))
        return fgpu_fail(FGPU_ERR_SCHEME, std::string("Unknown vertical advection type ") +
                                              (tra_adv_ver ? tra_adv_ver : "(null)") + "! Check your namelists!");
    if (parse_scheme(tra_adv_lim, lim_names, 2, &l))
        return fgpu_fail(FGPU_ERR_SCHEME, std::string("Unknown limiter ") +
                                              (tra_adv_lim ? tra_adv_lim : "(null)"));
    if (AB_order != 2 && AB_order != 3)
        return fgpu_fail(FGPU_ERR_ARG, "Adams-Bashfort tracer order must be 2 or 3");
    switch (ID) {   // bc_surface IDs of a plain build (oce_ale_tracer.F90:1862-2098)
    case 1: case 2: case 100: case 101: case 102: case 103: case 104: case 301: case 302: case 303: case 304:
    case 501: break;
    default:
        return fgpu_fail(FGPU_ERR_BC, "invalid ID " + std::to_string(ID) +
                                          " specified in boundary conditions (supported: 1, 2, 100-104, 301-304, 501)");
    }
    const int i = tr_num - 1;
    FGPU_HIP(hipSetDevice(c->device));
    DevTracers &t = c->t;
    if (AB_order == 3 && t.sOld < 2 * t.sN) {
        // valuesold(2,nl-1,nod2D) for every tracer: the first time level of each
        // tracer's old block is kept, the second starts at zero
        FGPU_HIP(hipStreamSynchronize(c->stream));
        double *grown = nullptr;
        const size_t nt = (size_t)t.ntr;
        int rc = fgpu_alloc(c, (void **)&grown, 2 * t.sN * nt * sizeof(double), true);
        if (rc) return rc;
        FGPU_HIP(hipMemset(grown, 0, 2 * t.sN * nt * sizeof(double)));
        for (size_t k = 0; k < nt; ++k)
            FGPU_HIP(hipMemcpy(grown + 2 * t.sN * k, t.Told + t.sOld * k, t.sOld * sizeof(double),
                               hipMemcpyDeviceToDevice));
        free_tracer_buf(c, t.Told);
        t.Told = grown;
        t.sOld = 2 * t.sN;
    }
    c->h_id[i] = ID; c->h_abo[i] = AB_order; c->h_hor[i] = h; c->h_ver[i] = v; c->h_lim[i] = l;
    c->h_ph[i] = tra_adv_ph; c->h_pv[i] = tra_adv_pv; c->h_ivd[i] = i_vert_diff ? 1 : 0;
    c->h_configured[i] = 1;
    c->graph_valid = false;
    return push_config(c);
}

extern "C" int fgpu_tracer_upload(fgpu_context *c, int tr_num, const double *values, const double *valuesold)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->tracers_ready) return fgpu_fail(FGPU_ERR_STATE, "tracers not initialised");
    if (tr_num < 1 || tr_num > c->t.ntr) return fgpu_fail(FGPU_ERR_ARG, "tr_num out of range");
    if (!values || !valuesold) return fgpu_fail(FGPU_ERR_ARG, "null array");
    FGPU_HIP(hipSetDevice(c->device));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    const int i = tr_num - 1;
    const size_t nold = c->t.sN * (size_t)(c->h_abo[i] - 1);
    FGPU_HIP(hipMemcpy(c->t.T + c->t.sN * i, values, c->t.sN * sizeof(double), hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(c->t.Told + c->t.sOld * i, valuesold, nold * sizeof(double), hipMemcpyHostToDevice));
    return FGPU_OK;
}

extern "C" int fgpu_tracer_download(fgpu_context *c, int tr_num, double *values, double *valuesold)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->tracers_ready) return fgpu_fail(FGPU_ERR_STATE, "tracers not initialised");
    if (tr_num < 1 || tr_num > c->t.ntr) return fgpu_fail(FGPU_ERR_ARG, "tr_num out of range");
    FGPU_HIP(hipSetDevice(c->device));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    const int i = tr_num - 1;
    const size_t nold = c->t.sN * (size_t)(c->h_abo[i] - 1);
    if (values)
        FGPU_HIP(hipMemcpy(values, c->t.T + c->t.sN * i, c->t.sN * sizeof(double), hipMemcpyDeviceToHost));
    if (valuesold)
        FGPU_HIP(hipMemcpy(valuesold, c->t.Told + c->t.sOld * i, nold * sizeof(double), hipMemcpyDeviceToHost));
    return FGPU_OK;
}

extern "C" int fgpu_tracer_work_download(fgpu_context *c, int tr_num, const char *name, double *out, long n)
{
    if (!c || !name || !out) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (!c->tracers_ready) return fgpu_fail(FGPU_ERR_STATE, "tracers not initialised");
    if (tr_num < 1 || tr_num > c->t.ntr) return fgpu_fail(FGPU_ERR_ARG, "tr_num out of range");
    FGPU_HIP(hipSetDevice(c->device));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    const size_t i = tr_num - 1;
    const DevTracers &t = c->t;
    const double *src = nullptr;
    size_t cnt = 0;
    std::string s(name);
    if (s == "valuesAB") { src = t.TAB ? t.TAB + t.sN * i : nullptr; cnt = t.sN; }
    else if (s == "tr_xy") { src = t.tr_xy ? t.tr_xy + t.sStg * i : nullptr; cnt = t.sF2; }
    else if (s == "edge_up_dn_grad") { src = t.dbg_grad; cnt = t.sE * 4; }
    else if (s == "fct_LO") { src = t.fct_lo + t.sN * i; cnt = t.sN; }
    else if (s == "adv_flux_hor") { src = t.a_h + t.sE * i; cnt = t.sE; }
    else if (s == "adv_flux_ver") { src = t.dbg_av; cnt = t.sNI; }
    else if (s == "del_ttf_advhoriz") { src = t.dttf_h; cnt = t.sN; }
    else if (s == "del_ttf_advvert") { src = t.dttf_v; cnt = t.sN; }
    else if (s == "fct_plus") { src = t.rp + t.sN * i; cnt = t.sN; }
    else if (s == "fct_minus") { src = t.rm + t.sN * i; cnt = t.sN; }
    else return fgpu_fail(FGPU_ERR_ARG, "unknown work array " + s);
    if (!src) return fgpu_fail(FGPU_ERR_STATE, "work array " + s + " exists after a per-tracer call only");
    if ((size_t)n != cnt) return fgpu_fail(FGPU_ERR_ARG, "work array " + s + " has " + std::to_string(cnt) + " entries");
    FGPU_HIP(hipMemcpy(out, src, cnt * sizeof(double), hipMemcpyDeviceToHost));
    return FGPU_OK;
}
