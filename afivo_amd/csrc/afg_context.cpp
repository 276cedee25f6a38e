// afg_context.cpp -- host side of the afivo_amd C ABI: tree tables, device
// pools, data movement and the orchestration of the FAS multigrid cycles.
//
// The cycle orchestration restates m_aX_mg.f90 (mg$D_fas_fmg :164-203,
// mg$D_fas_vcycle :208-255 and the stages they call) stage by stage; every
// stage is one or a few kernel launches over all boxes of a level.

#include "afivo_amd.h"
#include "afg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <map>
#include <sched.h>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

using namespace afg;

namespace {

thread_local std::string g_create_error;

struct Level {
    int lvl = 0, nc = 0, slot0 = 0, n = 0;
    double dr = 0, dx2 = 0, inv_dr_sq = 0;
    int parents_off = 0, n_parents = 0;
    int leaves_off = 0, n_leaves = 0;
    int all_off = 0;
    // In a forest the boxes of a level are ordered by their tree's highest
    // level, deepest trees first; n_act[b - min_lvl] boxes (parents, leaves)
    // belong to trees whose highest level is >= b (b = min_lvl..highest+1).
    std::vector<int> n_act, n_par_act, n_leaf_act;
    int tile[2] = {0, 0};   // tile shape of the colour pass: FP64, FP32 stages
    bool pairs[2] = {false, false};   // colour passes as fused pairs: FP64, FP32 stages
};

struct BcRec {
    int slot, face, type;
    int voff;   // -1: b = 0
};

// An instantiated V-cycle graph; in timing mode it also owns the event pairs
// recorded around its colour-pass launches and their algorithmic bytes.
// A timed block of launches: their number and summed algorithmic bytes.
struct TimedBlock {
    int n;
    double bytes;
};

struct GraphRec {
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<TimedBlock> blocks;
    bool pending = false;   // launched since its events were last read
};

void destroy_graph(GraphRec &g)
{
    if (g.exec) hipGraphExecDestroy(g.exec);
    for (auto &e : g.ev) {
        hipEventDestroy(e.first);
        hipEventDestroy(e.second);
    }
    g.exec = nullptr;
    g.ev.clear();
    g.blocks.clear();
}

uint64_t spread2(uint64_t x)   // 2D interleave (bit_space_1, m_morton.f90)
{
    x &= 0xffffffffULL;
    x = (x | (x << 16)) & 0x0000ffff0000ffffULL;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffULL;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

uint64_t spread3(uint64_t x)   // 3D interleave (bit_space_2, m_morton.f90)
{
    x &= 0x1fffffULL;
    x = (x | (x << 32)) & 0x001f00000000ffffULL;
    x = (x | (x << 16)) & 0x001f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

}  // namespace

struct afg_ctx {
    int dim = 0, nc = 0, n_var = 0, nch = 0, nnb = 0, stride = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // forest: trees solved together (one tree unless afg_set_forest says otherwise)
    int n_trees = 1;
    std::vector<int> tree_hi;            // highest level of each tree
    int min_tree_hi = 0;                 // the smallest of them
    int bound = 0;                       // stages take trees with highest level >= bound
    int res_upper = 1 << 20;             // ... and store residuals of trees below res_upper
    // leaf slots grouped by tree, cut into chunks of at most ~4096 cells (each
    // within one tree): chunk g = leaves [d_chunk[g], d_chunk[g+1]), tree t =
    // chunks [d_tree_chunk[t], d_tree_chunk[t+1])
    int32_t *d_tree_leaves = nullptr, *d_chunk = nullptr, *d_tree_chunk = nullptr;
    int n_chunks = 0;
    double *d_chunk_max = nullptr, *d_tree_out = nullptr, *h_tree_out = nullptr;
    // mixed-precision solve in steps: begun, and the next correction is the first (an FMG)
    bool mixed_ready = false, mixed_fmg = false;
    bool mixed_first_fmg = true;         // the first correction after afg_mixed_begin is an FMG
    bool mixed_base64 = false;           // FP32 stages: the coarsest-grid relaxation in FP64 arithmetic
    bool mixed_tau64 = false;            // FP32 stages: update_coarse (the FAS tau) in FP64 arithmetic
    bool mixed_r32 = false;              // the FP32 rhs holds (float) of the FP64 residual in tmp
    const void *mixed_r32_of[2] = {nullptr, nullptr};   // that tmp pool and that FP32 rhs pool
    bool tags_set = false;               // every box has its eps tag (they are set once)

    // tree
    bool have_tree = false;
    int min_lvl = 0, highest_lvl = -1, n_slots = 0, max_id = 0;
    double dr_base = 0;
    std::vector<Level> levels;           // index lvl - min_lvl
    std::vector<int32_t> id2slot;        // -1 if absent
    std::vector<int32_t> slot2id, slot_lvl, slot_nc;
    std::vector<int32_t> slot_ix, slot_tree;   // box%ix (3 per slot) and tree index per slot
    std::vector<int32_t> slot_par;             // parent slot, -1 if none
    bool rb_parent_nb = true;            // every refinement-boundary face has a neighbour of the parent there
    // cylindrical coordinates (2D): x is the radius, r_base the radius at ix = 1
    bool cyl = false;
    double r_base = 0;
    int32_t *d_cyl_ioff = nullptr;
    double *d_cyl_rmin = nullptr;
    std::vector<int32_t> h_nb;           // slot * nnb + face (afivo ids)
    std::vector<BcRec> bc;
    std::vector<int32_t> bc_of;          // slot * nnb + face -> bc record or -1
    std::vector<double> bc_vals;
    bool bc_dirty = false;

    // device
    int32_t *d_nb = nullptr, *d_parent = nullptr, *d_choff = nullptr, *d_lists = nullptr;
    int32_t *d_bc_type = nullptr, *d_bc_voff = nullptr;
    double *d_bc_vals = nullptr;
    size_t bc_vals_cap = 0, bc_cap = 0;
    int8_t *d_tag = nullptr;
    // bottom kernel tables, index 2 * fp32 + (eps in LDS): FP64 / FP32 values,
    // 3 / 4 variables in LDS
    BottomLevels *d_bottom[4] = {nullptr, nullptr, nullptr, nullptr};
    BottomLevels h_bottom[4];                         // their host copies
    BaseCell *d_base[4] = {nullptr, nullptr, nullptr, nullptr};   // lowest-level cell lists
    int base_cap = 0;
    // persistent kernel tables (the same four), its lowest-level cell list
    // (positions in the HBM pools) and its barrier counters
    BottomLevels *d_persist[4] = {nullptr, nullptr, nullptr, nullptr};
    BottomLevels h_persist[4];
    BaseCell *d_pbase = nullptr;
    int pbase_cap = 0;
    unsigned *d_psync = nullptr;
    unsigned *h_psync = nullptr;   // pinned: where a synchronisation reads them back
    unsigned long long *d_ptrace = nullptr;   // stage trace of the persistent kernel (afg_persist_trace)
    int ptrace_cap = 0;
    bool persist_ran = false;      // launched since the last read-back
    int8_t *d_is_parent = nullptr;
    std::vector<double *> d_var;
    std::vector<float *> d_var32;        // FP32 copies (mixed-precision solve), allocated on use
    // face variables (box%fx, fy, fz): fstride doubles per slot, allocated on use
    std::vector<double *> d_fc;
    int fstride = 0;
    // a$D_consistent_fluxes work records, by level (built on first use)
    bool flux_ready = false;
    std::vector<int> flux_off;
    FluxRec *d_flux = nullptr;
    // second phi pool of the fused pass pairs (FP64, FP32): a level's phi goes
    // there and back within a relaxation block
    double *d_phi2 = nullptr;
    float *d_phi2_32 = nullptr;
    // rhs by colour without ghosts (FP64, FP32 stages), read by the tiled
    // colour passes; valid while it equals the interiors of variable rb_var
    double *d_rhs_rb = nullptr;
    float *d_rhs_rb32 = nullptr;
    // phi by colour (FP64, FP32 stages): the working copy of the colour passes
    // within a relaxation block, valid only inside the block
    double *d_phi_rb = nullptr;
    float *d_phi_rb32 = nullptr;
    bool split_passes = true;            // relaxation blocks on the colour-split copy (3D tiles)
    bool rb_valid[2] = {false, false};
    int rb_var[2] = {-1, -1};
    double *d_partial = nullptr;
    int partial_cap = 0;
    double *d_out = nullptr, *h_out = nullptr;
    double *scratch = nullptr;           // this device's scratch array (afg_tree_extreme_noalloc)
    double *h_stage = nullptr;
    size_t stage_cap = 0;
    hipEvent_t xfer_ev[8] = {};          // per piece of a download (afg_download_boxes)

    // execution controls
    int bottom_lvl[4] = {-1000, -1000, -1000, -1000};  // highest level run by the bottom kernel
    size_t bottom_lds[4] = {0, 0, 0, 0};               // its dynamic LDS bytes
    int bottom_threads = BOTTOM_BLOCK;   // its workgroup size
    int64_t lds_limit = 65536;           // LDS bytes one workgroup may use
    int64_t bottom_cells[2] = {2048, 2048};   // cell limit per level for the bottom kernel: FP64, FP32
    int64_t bottom_wave_cells = 0;      // ... and for its one-wavefront levels
    // The persistent kernel (in place of the bottom kernel when persist_wgs >
    // 0): its workgroups, the XCD they run on, its cell limit per level (FP64,
    // FP32) and the highest level it takes per table (before the timing cap)
    int persist_wgs = 0;
    int persist_xcd = 0;
    int64_t persist_cells[2] = {32768, 32768};
    int persist_lvl[4] = {-1000, -1000, -1000, -1000};
    int64_t persist_max_cells[4] = {0, 0, 0, 0};      // most cells of one of its levels
    bool use_graphs = true;
    std::map<std::tuple<int, int, int, int, int, int, int, int, int, int, int>, GraphRec> graphs;
    bool capturing = false;
    bool timing = false;
    int timing_lvl = -1000;             // level whose passes are timed (below min_lvl: all)
    unsigned timing_ring = 0;           // replay counter: graph copy used in timing mode
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_used, ev_free;
    std::vector<TimedBlock> ev_blocks;
    // events recorded into the graph being captured (timing mode)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> cap_ev;
    std::vector<TimedBlock> cap_blocks;
    int64_t t_launches = 0;
    double t_ms = 0, t_bytes = 0;
    // every device allocation of the context and its size (afg_device_bytes)
    std::unordered_map<void *, size_t> dev_alloc;
    int64_t dev_bytes = 0;

    int fail(int code, const char *fmt, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

// Every device allocation of a context goes through these two, so that
// afg_device_bytes reports the context's whole footprint.
template <typename T>
hipError_t ctx_malloc(afg_ctx *c, T **p, size_t bytes)
{
    const hipError_t e = hipMalloc((void **)p, bytes);
    if (e == hipSuccess) {
        c->dev_alloc[(void *)*p] = bytes;
        c->dev_bytes += (int64_t)bytes;
    }
    return e;
}

void ctx_free(afg_ctx *c, void *p)
{
    if (!p) return;
    auto it = c->dev_alloc.find(p);
    if (it != c->dev_alloc.end()) {
        c->dev_bytes -= (int64_t)it->second;
        c->dev_alloc.erase(it);
    }
    hipFree(p);
}

// A kernel launch that failed (bad configuration, no code object for the
// device) leaves its error for hipGetLastError; the public entry points that
// launch kernels end with this check.
int launched(afg_ctx *c, const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return c->fail(AFG_ERR_HIP, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return 0;
}

#define HIP_TRY(ctx, expr)                                                            \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return (ctx)->fail(AFG_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Host-to-device copy on the context's own (non-blocking) stream, complete on
// return.  The legacy stream is never used: another thread may be capturing
// a graph on its own stream meanwhile.
#define HIP_H2D(ctx, dst, src, bytes)                                                       \
    do {                                                                                    \
        HIP_TRY(ctx, hipMemcpyAsync((dst), (src), (bytes), hipMemcpyHostToDevice, (ctx)->stream)); \
        HIP_TRY(ctx, hipStreamSynchronize((ctx)->stream));                                  \
    } while (0)


namespace {

int choose_bottom(afg_ctx *c);

// Synchronise the context's stream.  When the persistent kernel ran since
// the last synchronisation, its barrier counters are read back with it: a
// set error flag means a barrier wait timed out (its workgroups were not all
// resident at once), so the results of that work are invalid.  The counters
// are then cleared, the persistent kernel is turned off for the context and
// the call fails.
int stream_sync(afg_ctx *c)
{
    const bool check = c->persist_ran && !c->capturing;
    if (check)
        HIP_TRY(c, hipMemcpyAsync(c->h_psync, c->d_psync, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (!check) return 0;
    c->persist_ran = false;
    if (c->h_psync[2] == 0) return 0;
    HIP_TRY(c, hipMemsetAsync(c->d_psync, 0, 4 * sizeof(unsigned), c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const unsigned code = c->h_psync[2];
    c->persist_wgs = 0;
    choose_bottom(c);
    return c->fail(AFG_ERR_HIP, "persistent kernel: %s; the results since the last synchronisation are invalid and "
                   "the persistent kernel is now off",
                   code == 2 ? "its workgroups did not all run on one XCD" : "a barrier wait timed out");
}

DevTree dev_tree(const afg_ctx *c)
{
    DevTree T;
    T.stride = c->stride;
    T.nb = c->d_nb;
    T.parent = c->d_parent;
    T.choff = c->d_choff;
    T.bc_type = c->d_bc_type;
    T.bc_voff = c->d_bc_voff;
    T.bc_vals = c->d_bc_vals;
    T.tag = c->d_tag;
    T.is_parent = c->d_is_parent;
    T.cyl_ioff = c->d_cyl_ioff;
    T.cyl_rmin = c->d_cyl_rmin;
    return T;
}

LevelArgs level_args(const afg_ctx *c, int lvl)
{
    const Level &L = c->levels[lvl - c->min_lvl];
    LevelArgs a;
    a.slot0 = L.slot0;
    a.n = L.n_act.empty() ? L.n : L.n_act[c->bound - c->min_lvl];
    a.nc = L.nc;
    a.tile = 0;
    a.pnc = lvl > c->min_lvl ? c->levels[lvl - 1 - c->min_lvl].nc : 0;
    a.dr = L.dr;
    a.dx2 = L.dx2;
    a.inv_dr_sq = L.inv_dr_sq;
    return a;
}

ListArgs list_args(const afg_ctx *c, int lvl, int which)   // 0 parents, 1 leaves, 2 all
{
    const Level &L = c->levels[lvl - c->min_lvl];
    ListArgs a;
    const int off = which == 0 ? L.parents_off : which == 1 ? L.leaves_off : L.all_off;
    a.slots = c->d_lists + off;
    const int b = c->bound - c->min_lvl;
    if (L.n_act.empty())
        a.n = which == 0 ? L.n_parents : which == 1 ? L.n_leaves : L.n;
    else
        a.n = which == 0 ? L.n_par_act[b] : which == 1 ? L.n_leaf_act[b] : L.n_act[b];
    a.nc = L.nc;
    a.dr = L.dr;
    a.dx2 = L.dx2;
    a.inv_dr_sq = L.inv_dr_sq;
    return a;
}

int box_len(int dim, int n) { return dim == 2 ? (n + 2) * (n + 2) : (n + 2) * (n + 2) * (n + 2); }

void free_tree(afg_ctx *c)
{
    for (auto &g : c->graphs) destroy_graph(g.second);
    c->graphs.clear();
    ctx_free(c, c->d_nb);
    ctx_free(c, c->d_parent);
    ctx_free(c, c->d_choff);
    ctx_free(c, c->d_lists);
    ctx_free(c, c->d_bc_type);
    ctx_free(c, c->d_bc_voff);
    ctx_free(c, c->d_bc_vals);
    ctx_free(c, c->d_tag);
    ctx_free(c, c->d_is_parent);
    c->d_is_parent = nullptr;
    for (int k = 0; k < 4; k++) {
        ctx_free(c, c->d_bottom[k]);
        ctx_free(c, c->d_base[k]);
        ctx_free(c, c->d_persist[k]);
        c->d_bottom[k] = nullptr;
        c->d_base[k] = nullptr;
        c->d_persist[k] = nullptr;
    }
    c->base_cap = 0;
    ctx_free(c, c->d_pbase);
    c->d_pbase = nullptr;
    c->pbase_cap = 0;
    ctx_free(c, c->d_tree_leaves);
    ctx_free(c, c->d_chunk);
    ctx_free(c, c->d_tree_chunk);
    ctx_free(c, c->d_chunk_max);
    ctx_free(c, c->d_tree_out);
    if (c->h_tree_out) hipHostFree(c->h_tree_out);
    c->d_tree_leaves = c->d_chunk = c->d_tree_chunk = nullptr;
    c->d_chunk_max = c->d_tree_out = c->h_tree_out = nullptr;
    c->n_chunks = 0;
    ctx_free(c, c->d_partial);
    for (double *p : c->d_var) ctx_free(c, p);
    for (float *p : c->d_var32) ctx_free(c, p);
    for (double *p : c->d_fc) ctx_free(c, p);
    std::fill(c->d_fc.begin(), c->d_fc.end(), nullptr);
    ctx_free(c, c->d_flux);
    c->d_flux = nullptr;
    ctx_free(c, c->d_cyl_ioff);
    ctx_free(c, c->d_cyl_rmin);
    c->d_cyl_ioff = nullptr;
    c->d_cyl_rmin = nullptr;
    c->flux_ready = false;
    ctx_free(c, c->d_phi2);
    ctx_free(c, c->d_phi2_32);
    c->d_phi2 = nullptr;
    c->d_phi2_32 = nullptr;
    ctx_free(c, c->d_rhs_rb);
    ctx_free(c, c->d_rhs_rb32);
    c->d_rhs_rb = nullptr;
    c->d_rhs_rb32 = nullptr;
    ctx_free(c, c->d_phi_rb);
    ctx_free(c, c->d_phi_rb32);
    c->d_phi_rb = nullptr;
    c->d_phi_rb32 = nullptr;
    c->rb_valid[0] = c->rb_valid[1] = false;
    c->mixed_r32 = false;
    c->d_nb = c->d_parent = c->d_choff = c->d_lists = c->d_bc_type = c->d_bc_voff = nullptr;
    c->d_bc_vals = nullptr;
    c->d_tag = nullptr;
    c->d_partial = nullptr;
    c->partial_cap = 0;
    c->bc_vals_cap = c->bc_cap = 0;
    std::fill(c->d_var.begin(), c->d_var.end(), nullptr);
    std::fill(c->d_var32.begin(), c->d_var32.end(), nullptr);
    c->have_tree = false;
}

int check_ctx(afg_ctx *c)
{
    if (!c) return AFG_ERR_ARG;
    if (!c->have_tree) return c->fail(AFG_ERR_STATE, "no tree set (call afg_set_tree first)");
    HIP_TRY(c, hipSetDevice(c->device));
    (void)hipGetLastError();   // a launch error reported later is one of this call's
    return 0;
}

int check_var(afg_ctx *c, int iv)
{
    if (iv < 1 || iv > c->n_var) return c->fail(AFG_ERR_ARG, "variable index %d out of range 1..%d", iv, c->n_var);
    return 0;
}

// The device pool of variable iv (FP64, or its FP32 copy), allocated and
// zeroed on first use.
template <typename R>
int ensure_pool(afg_ctx *c, int iv, std::vector<R *> &pools, R **out)
{
    if (int e = check_var(c, iv)) return e;
    R *&p = pools[iv - 1];
    if (!p) {
        if (c->capturing) return c->fail(AFG_ERR_STATE, "variable %d first used inside a graph capture", iv);
        const size_t bytes = (size_t)c->n_slots * c->stride * sizeof(R);
        hipError_t e = ctx_malloc(c, &p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return c->fail(AFG_ERR_MEMORY, "hipMalloc(%zu) for variable %d: %s", bytes, iv,
                           hipGetErrorString(e));
        }
        HIP_TRY(c, hipMemsetAsync(p, 0, bytes, c->stream));
    }
    if (out) *out = p;
    return 0;
}

int ensure_var(afg_ctx *c, int iv, double **out) { return ensure_pool(c, iv, c->d_var, out); }
int ensure_var(afg_ctx *c, int iv, float **out) { return ensure_pool(c, iv, c->d_var32, out); }

// The device pool of face variable ifc (1..AFG_MAX_FC), allocated and zeroed
// on first use.
int ensure_fc(afg_ctx *c, int ifc, double **out)
{
    if (ifc < 1 || ifc > AFG_MAX_FC) return c->fail(AFG_ERR_ARG, "face variable %d out of range 1..%d", ifc, AFG_MAX_FC);
    double *&p = c->d_fc[ifc - 1];
    if (!p) {
        if (c->capturing) return c->fail(AFG_ERR_STATE, "face variable %d first used inside a graph capture", ifc);
        const size_t bytes = (size_t)c->n_slots * c->fstride * sizeof(double);
        if (ctx_malloc(c, &p, bytes) != hipSuccess) {
            p = nullptr;
            return c->fail(AFG_ERR_MEMORY, "hipMalloc(%zu) for face variable %d", bytes, ifc);
        }
        HIP_TRY(c, hipMemsetAsync(p, 0, bytes, c->stream));
    }
    *out = p;
    return 0;
}

int build_base(afg_ctx *c);
int choose_persist(afg_ctx *c);

int upload_bc(afg_ctx *c)
{
    if (!c->bc_dirty) return 0;
    const size_t nrec = std::max<size_t>(c->bc.size(), 1);
    std::vector<int32_t> ty(nrec, -1), vo(nrec, -1);
    for (size_t r = 0; r < c->bc.size(); r++) {
        ty[r] = c->bc[r].type;
        vo[r] = c->bc[r].voff;
    }
    if (nrec > c->bc_cap) {
        ctx_free(c, c->d_bc_type);
        ctx_free(c, c->d_bc_voff);
        c->d_bc_type = c->d_bc_voff = nullptr;
        HIP_TRY(c, ctx_malloc(c, &c->d_bc_type, nrec * sizeof(int32_t)));
        HIP_TRY(c, ctx_malloc(c, &c->d_bc_voff, nrec * sizeof(int32_t)));
        c->bc_cap = nrec;
        for (auto &g : c->graphs) destroy_graph(g.second);
        c->graphs.clear();
    }
    const size_t nv = std::max<size_t>(c->bc_vals.size(), 1);
    if (nv > c->bc_vals_cap) {
        ctx_free(c, c->d_bc_vals);
        c->d_bc_vals = nullptr;
        HIP_TRY(c, ctx_malloc(c, &c->d_bc_vals, nv * sizeof(double)));
        c->bc_vals_cap = nv;
        for (auto &g : c->graphs) destroy_graph(g.second);
        c->graphs.clear();
    }
    HIP_TRY(c, hipMemcpyAsync(c->d_bc_type, ty.data(), nrec * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_bc_voff, vo.data(), nrec * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    if (!c->bc_vals.empty())
        HIP_TRY(c, hipMemcpyAsync(c->d_bc_vals, c->bc_vals.data(), c->bc_vals.size() * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream));
    if (int e_ = stream_sync(c)) return e_;
    c->bc_dirty = false;
    return c->have_tree ? build_base(c) : 0;   // the boundary values enter the base cell lists
}

int ensure_stage(afg_ctx *c, size_t n_doubles)
{
    if (n_doubles <= c->stage_cap) return 0;
    if (c->h_stage) hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    HIP_TRY(c, hipHostMalloc((void **)&c->h_stage, n_doubles * sizeof(double), hipHostMallocDefault));
    c->stage_cap = n_doubles;
    return 0;
}

// ---- timing of the colour-pass kernel ----

// Algorithmic bytes of one colour pass over a level: half of a full sweep's
// (3 + 2*dim/nc) * 8 per cell (read phi with its face ghosts, read rhs, write
// phi), plus (1 + 2*dim/nc) * 8 per cell of eps reads for the auto operators.
template <typename R>
double gsrb_pass_bytes(const afg_ctx *c, const LevelArgs &L, const MgVarsT<R> &V)
{
    const double cells = (double)L.n * (c->dim == 2 ? L.nc * L.nc : L.nc * L.nc * L.nc);
    const double g = 2.0 * c->dim / L.nc;
    double per_cell = (3 + g) * sizeof(R);
    if (V.op_auto) per_cell += (1 + g) * sizeof(R);
    return 0.5 * cells * per_cell;
}

// Algorithmic bytes of the ghost fill of a level: a value read and a ghost
// written per face ghost cell (2D/nc ghost cells per cell).
template <typename R>
double fill_bytes(const afg_ctx *c, const LevelArgs &L)
{
    const double ghosts = (double)L.n * 2 * c->dim * (c->dim == 2 ? L.nc : L.nc * L.nc);
    return ghosts * 2 * sizeof(R);
}

// Consecutive colour passes of one level, bracketed by one pair of HIP
// events on the context stream when that level is timed (a timed block holds
// n launches and their summed algorithmic bytes).  Where the passes run on
// the colour-split copy of phi: with `fill` the last pass also does the
// level's ghost fill (and the parents' diff with `diff`); with `keep` the
// block leaves phi in the copy (the box arrays keep the state before it) and
// *in_copy says so; with `from_copy` the block starts from the copy (the
// state the block before left there, as corrected since).  Returns whether
// the fill was done.
template <typename R>
bool timed_passes(afg_ctx *c, const DevTree &T, const LevelArgs &L, int lvl, const MgVarsT<R> &V, R *phi2,
                  int n_pass, int first_fly, int fill = 0, const ListArgs *diff = nullptr, int keep = 0,
                  int from_copy = 0, bool *in_copy = nullptr)
{
    const bool on = c->timing && n_pass > 0 && (c->timing_lvl < c->min_lvl || lvl == c->timing_lvl);
    std::pair<hipEvent_t, hipEvent_t> ev;
    // inside a capture the events must become event-record nodes of the graph
    const unsigned flags = c->capturing ? hipEventRecordExternal : 0;
    if (on) {
        if (!c->capturing && !c->ev_free.empty()) {
            ev = c->ev_free.back();
            c->ev_free.pop_back();
        } else {
            hipEventCreate(&ev.first);
            hipEventCreate(&ev.second);
        }
        hipEventRecordWithFlags(ev.first, c->stream, flags);
    }
    LevelArgs Lt = L;
    const Level &Lv = c->levels[lvl - c->min_lvl];
    Lt.tile = Lv.tile[sizeof(R) == 4];
    // Passes go by fused pairs, phi to the second pool and back, while an
    // even number of pairs is left (on levels set to pairs); the rest one by
    // one: on the colour-split copy of phi when two or more are left (first,
    // middle and last passes), else on the box arrays.
    int rb = 1, launches = 0;
    if (in_copy) *in_copy = false;
    if (phi2 && Lv.pairs[sizeof(R) == 4] && !from_copy)
        for (; rb + 3 <= n_pass; rb += 4) {
            if (!launch_gsrb_pair(c->dim, c->stream, T, Lt, V, V.phi, phi2, rb, rb == 1 ? first_fly : 1)) break;
            launch_gsrb_pair(c->dim, c->stream, T, Lt, V, phi2, V.phi, rb + 2, 1);
            launches += 2;
        }
    const int first = rb;
    const bool split = n_pass - first >= 1 && launch_gsrb_split(c->dim, c->stream, T, Lt, V, rb, from_copy ? 1 : 0,
                                                                 rb == 1 ? first_fly : 1);
    if (split) {
        for (rb++, launches++; rb <= n_pass; rb++, launches++)
            launch_gsrb_split(c->dim, c->stream, T, Lt, V, rb, rb < n_pass ? 1 : fill ? 3 : keep ? 1 : 2, 1,
                              rb < n_pass ? nullptr : diff);
        if (in_copy) *in_copy = keep && !fill;
    }
    for (; rb <= n_pass; rb++, launches++)
        launch_gsrb_pass(c->dim, c->stream, T, Lt, V, rb, rb == 1 ? first_fly : 1);
    const bool filled = split && fill;
    if (!on) return filled;
    hipEventRecordWithFlags(ev.second, c->stream, flags);
    const double pb = diff ? (double)(diff->nc + 2) * (diff->nc + 2) * (c->dim == 3 ? diff->nc + 2 : 1) : 0.0;
    const double diff_bytes = filled && diff ? 3.0 * sizeof(R) * diff->n * pb : 0.0;
    const TimedBlock tb{launches, n_pass * gsrb_pass_bytes(c, L, V) + (filled ? fill_bytes<R>(c, L) : 0.0) + diff_bytes};
    if (c->capturing) {
        c->cap_ev.push_back(ev);
        c->cap_blocks.push_back(tb);
    } else {
        c->ev_used.push_back(ev);
        c->ev_blocks.push_back(tb);
    }
    return filled;
}

// Add the event times of a graph's last replay to the totals.
int harvest_graph(afg_ctx *c, GraphRec &g)
{
    if (!g.pending) return 0;
    HIP_TRY(c, hipEventSynchronize(g.ev.back().second));
    for (size_t k = 0; k < g.ev.size(); k++) {
        float ms = 0;
        HIP_TRY(c, hipEventElapsedTime(&ms, g.ev[k].first, g.ev[k].second));
        c->t_ms += ms;
        c->t_bytes += g.blocks[k].bytes;
        c->t_launches += g.blocks[k].n;
    }
    g.pending = false;
    return 0;
}

int collect_timing(afg_ctx *c)
{
    for (auto &g : c->graphs)
        if (int e = harvest_graph(c, g.second)) return e;
    if (c->ev_used.empty()) return 0;
    if (int e_ = stream_sync(c)) return e_;
    for (size_t k = 0; k < c->ev_used.size(); k++) {
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev_used[k].first, c->ev_used[k].second);
        c->t_ms += ms;
        c->t_bytes += c->ev_blocks[k].bytes;
        c->t_launches += c->ev_blocks[k].n;
        c->ev_free.push_back(c->ev_used[k]);
    }
    c->ev_used.clear();
    c->ev_blocks.clear();
    return 0;
}

// ---- stages (m_aX_mg.f90) ----

// A solve's state: the context, its device tables, the variables (R = double:
// the cell variables; R = float: their FP32 copies) and the options.
template <typename R>
struct MgT {
    afg_ctx *c;
    DevTree T;
    MgVarsT<R> V;
    const afg_mg_opts *o;
    R *phi2 = nullptr;   // second phi pool of the pass pairs (null: single passes)
};
using Mg = MgT<double>;

double *&phi2_pool(afg_ctx *c, double *) { return c->d_phi2; }
float *&phi2_pool(afg_ctx *c, float *) { return c->d_phi2_32; }
double *&rhs_rb_pool(afg_ctx *c, double *) { return c->d_rhs_rb; }
float *&rhs_rb_pool(afg_ctx *c, float *) { return c->d_rhs_rb32; }
double *&phi_rb_pool(afg_ctx *c, double *) { return c->d_phi_rb; }
float *&phi_rb_pool(afg_ctx *c, float *) { return c->d_phi_rb32; }

// values per slot of the colour-split rhs: the interior of the largest boxes
int rb_stride(const afg_ctx *c) { return c->dim == 2 ? c->nc * c->nc : c->nc * c->nc * c->nc; }

// Cell data changed outside the multigrid stages: the colour-split rhs must
// be formed again before the next colour pass reads it.
void data_written(afg_ctx *c)
{
    c->rb_valid[0] = c->rb_valid[1] = false;
    c->mixed_r32 = false;
}

// Form the colour-split rhs of every level from rhs (consecutive levels of
// equal cell count in one launch) and mark it valid for variable iv.
template <typename R>
void refresh_rhs_rb(afg_ctx *c, const DevTree &T, const R *rhs, R *rb, int iv)
{
    const int nl = (int)c->levels.size();
    for (int a = 0; a < nl;) {
        LevelArgs L = level_args(c, c->levels[a].lvl);
        int b = a + 1;
        for (; b < nl && c->levels[b].nc == c->levels[a].nc; b++) L.n += c->levels[b].n;
        launch_split_rhs(c->dim, c->stream, T, L, rhs, rb, rb_stride(c));
        a = b;
    }
    c->rb_valid[sizeof(R) == 4] = true;
    c->rb_var[sizeof(R) == 4] = iv;
}

template <typename R>
int make_mg(afg_ctx *c, const afg_mg_opts *o, MgT<R> *m)
{
    if (!o) return c->fail(AFG_ERR_ARG, "null options");
    if (!o->initialized) return c->fail(AFG_ERR_STATE, "check_mg: you haven't called mg%d_init_mg", c->dim);
    if (o->op != AFG_OP_LPL && o->op != AFG_OP_AUTO && o->op != AFG_OP_LSF)
        return c->fail(AFG_ERR_UNSUPPORTED, "operator %d not supported on the device", o->op);
    if (o->rstr != AFG_RSTR_LPL && o->rstr != AFG_RSTR_AUTO && o->rstr != AFG_RSTR_LSF)
        return c->fail(AFG_ERR_UNSUPPORTED, "restriction %d not supported on the device", o->rstr);
    if (c->cyl && sizeof(R) == 4) return c->fail(AFG_ERR_UNSUPPORTED, "FP32 stages on a cylindrical tree");
    // the level set takes part with the lpllsf set, the auto set (its tags)
    // and the level-set restrictions
    const bool use_lsf = o->i_lsf > 0 && (o->op != AFG_OP_LPL || o->rstr != AFG_RSTR_LPL);
    if ((o->op == AFG_OP_LSF || o->rstr == AFG_RSTR_LSF) && o->i_lsf <= 0)
        return c->fail(AFG_ERR_ARG, "mg%d: the lpllsf operators need i_lsf and i_bval", c->dim);
    if (use_lsf && sizeof(R) == 4) return c->fail(AFG_ERR_UNSUPPORTED, "FP32 stages with the level-set operators");
    m->c = c;
    m->o = o;
    if (int e = ensure_var(c, o->i_phi, &m->V.phi)) return e;
    if (int e = ensure_var(c, o->i_rhs, &m->V.rhs)) return e;
    if (int e = ensure_var(c, o->i_tmp, &m->V.tmp)) return e;
    m->V.eps = nullptr;
    // the auto operators without eps leave every box untagged: the plain
    // operators, except on a cylindrical tree (mg2_box_clpl, examples/poisson_cyl.f90)
    m->V.op_auto = o->op == AFG_OP_AUTO && (o->i_eps != -1 || c->cyl || use_lsf);
    m->V.tau64 = sizeof(R) == 4 && c->mixed_tau64 ? 1 : 0;
    if (m->V.op_auto && o->i_eps != -1) {
        R *e = nullptr;
        if (int r = ensure_var(c, o->i_eps, &e)) return r;
        m->V.eps = e;
    }
    m->V.lsf = m->V.bval = nullptr;
    m->V.lsf_mode = m->V.rstr_mode = 0;
    if (use_lsf) {
        R *l = nullptr, *b = nullptr;
        if (int r = ensure_var(c, o->i_lsf, &l)) return r;
        if (int r = ensure_var(c, o->i_bval, &b)) return r;
        m->V.lsf = l;
        m->V.bval = b;
        m->V.lsf_mode = o->op == AFG_OP_LSF ? 2 : o->op == AFG_OP_AUTO ? 1 : 0;
        m->V.rstr_mode = o->rstr == AFG_RSTR_LSF ? 2 : o->rstr == AFG_RSTR_AUTO ? 1 : 0;
    }
    {
        R *&p2 = phi2_pool(c, (R *)nullptr);
        if (!p2) {
            if (c->capturing) return c->fail(AFG_ERR_STATE, "second phi pool first used inside a graph capture");
            const size_t bytes = (size_t)c->n_slots * c->stride * sizeof(R);
            if (ctx_malloc(c, &p2, bytes) != hipSuccess) {
                p2 = nullptr;
                return c->fail(AFG_ERR_MEMORY, "hipMalloc(%zu) for the second phi pool", bytes);
            }
            HIP_TRY(c, hipMemsetAsync(p2, 0, bytes, c->stream));
        }
        m->phi2 = p2;
    }
    if (int e = upload_bc(c)) return e;
    m->T = dev_tree(c);
    // the colour-split rhs, for the tiled colour passes in 3D (not on the
    // trees and operators that take the per-cell kernels only).  In 2D a box's
    // rhs rows are short and the fused pass pairs read the neighbours' rows
    // anyway, so the copy saves little there and the 2D kernels keep one rhs
    // path (fewer registers).
    m->V.rhs_rb = nullptr;
    m->V.rb_stride = rb_stride(c);
    if (c->dim == 3 && !per_cell_only(m->T, m->V) && !c->capturing) {
        const int k = sizeof(R) == 4;
        R *&rb = rhs_rb_pool(c, (R *)nullptr);
        if (!rb) {
            const size_t bytes = (size_t)c->n_slots * rb_stride(c) * sizeof(R);
            if (ctx_malloc(c, &rb, bytes) != hipSuccess) {
                rb = nullptr;
                return c->fail(AFG_ERR_MEMORY, "hipMalloc(%zu) for the colour-split rhs", bytes);
            }
            c->rb_valid[k] = false;
        }
        if (!c->rb_valid[k] || c->rb_var[k] != o->i_rhs) refresh_rhs_rb(c, m->T, m->V.rhs, rb, o->i_rhs);
        m->V.rhs_rb = rb;
        // the working copy of phi for the colour passes of a relaxation block
        R *&pr = phi_rb_pool(c, (R *)nullptr);
        if (!pr) {
            const size_t bytes = (size_t)c->n_slots * rb_stride(c) * sizeof(R);
            if (ctx_malloc(c, &pr, bytes) != hipSuccess) {
                pr = nullptr;
                return c->fail(AFG_ERR_MEMORY, "hipMalloc(%zu) for the colour-split phi", bytes);
            }
        }
        m->V.phi_rb = c->split_passes ? pr : nullptr;
    }
    return 0;
}

// The lazy tags of mg$D_set_box_tag (m_aX_mg.f90:667-706): set once per box,
// from the level set and eps at the first call of an auto operator (or of
// mg$D_auto_rstr, which needs them set), and never re-evaluated.  The first
// solve sets every box's; later calls have nothing to do.  With both a level
// set and eps the tags are read back once to report a box that has both
// (where the reference stops).
int set_tags(Mg &m)
{
    afg_ctx *c = m.c;
    if (!(m.V.op_auto || m.V.rstr_mode == 1) || c->tags_set) return 0;
    for (int l = c->min_lvl; l <= c->highest_lvl; l++)
        launch_set_tags(c->dim, c->stream, m.T, level_args(c, l), m.V.eps, m.V.lsf);
    c->tags_set = true;
    if (m.V.eps && m.V.lsf) {
        std::vector<int8_t> h(c->n_slots);
        HIP_TRY(c, hipMemcpyAsync(h.data(), m.T.tag, h.size(), hipMemcpyDeviceToHost, c->stream));
        if (int e_ = stream_sync(c)) return e_;
        for (int8_t t : h)
            if (t == TAG_CONFLICT)
                return c->fail(AFG_ERR_ARG, "mg%d_set_box_tag: Cannot set lsf and eps tag for same box", c->dim);
    }
    return 0;
}

template <typename R>
void fill_gc(MgT<R> &m, R *var, int lvl)
{
    launch_fill_gc(m.c->dim, m.c->stream, m.T, level_args(m.c, lvl), var);
}

// The colour passes of gsrb_boxes (m_aX_mg.f90:397-420) on one level.  The
// reference fills the ghost layer after every pass; nc is even on every
// level, so passes 2.. compute the face ghosts they read on the fly (the
// values that fill would store) and only the ghost fill after the last pass
// remains, which the caller launches or fuses into the next stage.  The
// first pass reads the ghosts in memory (first_fly = 0), as the reference
// does, or computes them when a ghost fill precedes it (first_fly = 1).
template <typename R>
int bottom_top(const MgT<R> &m);

// The variables for the stages of level lvl that read rhs: the colour-split
// copy only above the bottom kernel's levels (that kernel writes the rhs of
// its levels without the copy; the stages that write rhs keep it current on
// every level).
template <typename R>
MgVarsT<R> level_vars(const MgT<R> &m, int lvl)
{
    MgVarsT<R> V = m.V;
    if (lvl <= bottom_top(m)) V.rhs_rb = nullptr;
    return V;
}

template <typename R>
bool passes(MgT<R> &m, int lvl, int n_cycle, int first_fly, int fill = 0, const ListArgs *diff = nullptr,
            int keep = 0, int from_copy = 0, bool *in_copy = nullptr)
{
    const LevelArgs L = level_args(m.c, lvl);
    return timed_passes(m.c, m.T, L, lvl, level_vars(m, lvl), m.phi2, 2 * n_cycle, first_fly, fill, diff, keep,
                        from_copy, in_copy);
}

// gsrb_boxes as the reference has it: passes, then the last ghost fill.
template <typename R>
void gsrb_level(MgT<R> &m, int lvl, int n_cycle)
{
    passes(m, lvl, n_cycle, 0);
    if (n_cycle > 0) fill_gc(m, m.V.phi, lvl);
}

// update_coarse (m_aX_mg.f90:424-480) after `n_cycle` relaxations of lvl: the
// child stage reads lvl's face ghosts as the fill after the last relaxation
// stores them (that fill is not launched: nothing else reads those ghosts
// before the upward sweep fills lvl again), and the ghost fill of lvl-1 runs
// inside the parent stage.
template <typename R>
void update_coarse(MgT<R> &m, int lvl, int relaxed, bool in_copy = false)
{
    launch_uc_child(m.c->dim, m.c->stream, m.T, level_args(m.c, lvl), level_vars(m, lvl), 1, relaxed ? 1 : 0,
                    in_copy ? 1 : 0);
    launch_uc_parent(m.c->dim, m.c->stream, m.T, level_args(m.c, lvl - 1), m.V, 1);
}

// set_coarse_phi_rhs (m_aX_mg.f90:484-517)
template <typename R>
void set_coarse_phi_rhs(MgT<R> &m, int lvl)
{
    if (lvl == m.c->highest_lvl) fill_gc(m, m.V.phi, lvl);
    launch_uc_child(m.c->dim, m.c->stream, m.T, level_args(m.c, lvl), level_vars(m, lvl), 0, 0);
    launch_uc_parent(m.c->dim, m.c->stream, m.T, level_args(m.c, lvl - 1), m.V, 0);
}

// init_phi_rhs (m_aX_mg.f90:520-543)
template <typename R>
void init_phi_rhs(MgT<R> &m)
{
    for (int lvl = m.c->highest_lvl; lvl >= m.c->min_lvl + 1; lvl--) {
        const LevelArgs L = level_args(m.c, lvl);
        launch_fill_value(m.c->dim, m.c->stream, m.T, L.slot0, L.n, L.nc, m.V.phi, 0.0);
        launch_restrict(m.c->dim, m.c->stream, m.T, L, m.V.rhs, m.V.rhs, (const R *)m.V.lsf, m.V.rstr_mode);
    }
    if (m.V.rhs_rb) refresh_rhs_rb(m.c, m.T, (const R *)m.V.rhs, m.V.rhs_rb, m.o->i_rhs);
}

// correct_children over the parents of lvl-1 (m_aX_mg.f90:368-395) followed
// by relaxation and the ghost fill of lvl: the correction is formed on the
// fly from the parents' phi and tmp, and the parents' tmp = phi - tmp is
// written with the final ghost fill of lvl (nothing reads it in between).
template <typename R>
void correct_relax(MgT<R> &m, int lvl, int n_cycle, bool in_copy = false)
{
    // a level whose down-sweep block left phi in the colour-split copy is
    // corrected there and its up-sweep block starts from it
    launch_correct(m.c->dim, m.c->stream, m.T, level_args(m.c, lvl), level_vars(m, lvl), 1, in_copy ? 1 : 0);
    const ListArgs P = list_args(m.c, lvl - 1, 0);
    // the last pass on the colour-split copy fills the ghosts and forms the parents' tmp
    if (!passes(m, lvl, n_cycle, 1, 1, &P, 0, in_copy ? 1 : 0))
        launch_fill_diff(m.c->dim, m.c->stream, m.T, level_args(m.c, lvl), P, m.V);
}

// residual_box over level lvl, for the trees whose residual is stored: in a
// forest those with highest level below res_upper (a tail of the level's
// active boxes, deepest trees coming first)
template <typename R>
void residual_level(MgT<R> &m, int lvl)
{
    afg_ctx *c = m.c;
    LevelArgs a = level_args(c, lvl);
    const Level &L = c->levels[lvl - c->min_lvl];
    if (!L.n_act.empty() && c->res_upper <= c->highest_lvl) {
        const int skip = L.n_act[c->res_upper - c->min_lvl];
        a.slot0 += skip;
        a.n -= skip;
    }
    launch_residual(c->dim, c->stream, m.T, a, level_vars(m, lvl));
}

template <typename R>
BottomArgs bottom_args(MgT<R> &m, int top)
{
    afg_ctx *c = m.c;
    BottomArgs B;
    B.hl = top - c->min_lvl;
    B.n_down = m.o->n_cycle_down;
    B.n_up = m.o->n_cycle_up;
    B.n_base = m.o->n_cycle_base;
    B.base64 = sizeof(R) == 4 && c->mixed_base64 ? 1 : 0;
    B.tau64 = sizeof(R) == 4 && c->mixed_tau64 ? 1 : 0;
    B.trace = c->d_ptrace;
    B.trace_cap = c->ptrace_cap;
    return B;
}

// index of the bottom kernel table for these values and options
template <typename R>
int bottom_table(const MgT<R> &m) { return (sizeof(R) == sizeof(float) ? 2 : 0) + (m.V.op_auto ? 1 : 0); }

// highest level run by the bottom kernel -- or, when it is on, by the
// persistent kernel -- for these options.  In timing mode the timed level's
// passes run by level (their launches carry the events).
template <typename R>
int bottom_top(const MgT<R> &m)
{
    const afg_ctx *c = m.c;
    if (m.V.lsf) return -1000;
    if (c->persist_wgs <= 0) return c->bottom_lvl[bottom_table(m)];
    int top = c->persist_lvl[bottom_table(m)];
    if (c->timing && c->timing_lvl >= c->min_lvl && top >= c->timing_lvl) top = c->timing_lvl - 1;
    return top >= c->min_lvl ? top : -1000;
}


// mg$D_fas_vcycle (m_aX_mg.f90:208-255).  Levels up to bottom_top run in the
// one-workgroup LDS bottom kernel, or in the persistent kernel; the levels
// above, stage by stage.
template <typename R>
void vcycle_body(MgT<R> &m, int highest_lvl, int set_residual)
{
    afg_ctx *c = m.c;
    // the bottom and persistent kernels' tables cover every tree: used when all take part
    const int lb = c->bound <= c->min_tree_hi ? std::min(bottom_top(m), highest_lvl) : -1000;
    const int low = lb >= c->min_lvl ? lb : c->min_lvl;
    const bool persist = c->persist_wgs > 0;
    const int k = bottom_table(m);
    // the residual of the lower levels inside the persistent kernel: only when
    // it holds the whole V-cycle (above it, the up-sweep still writes tmp on
    // the parents of level lb + 1)
    const bool res_inside = persist && set_residual && lb == highest_lvl && c->res_upper > c->highest_lvl;
    // a level's phi may stay in the colour-split copy from its down-sweep block to
    // its up-sweep block (update_coarse and the correction work there), when an
    // up-sweep block will write it back; the FP64-tau stages read the box arrays
    const bool keep = m.o->n_cycle_up > 0 && !m.V.tau64;
    std::vector<char> in_copy(c->levels.size(), 0);
    for (int lvl = highest_lvl; lvl > low; lvl--) {
        bool ic = false;
        passes(m, lvl, m.o->n_cycle_down, 0, 0, nullptr, keep ? 1 : 0, 0, &ic);
        update_coarse(m, lvl, m.o->n_cycle_down > 0, ic);
        in_copy[lvl - c->min_lvl] = ic;
    }
    if (lb >= c->min_lvl && persist) {
        PersistArgs P;
        P.A = bottom_args(m, lb);
        P.set_residual = res_inside ? 1 : 0;
        P.n_rank = c->persist_wgs;
        P.xcd = c->persist_xcd;
        P.sync = c->d_psync;
        P.trace = c->d_ptrace;
        P.trace_cap = c->ptrace_cap;
        launch_persist(c->dim, c->stream, m.T, m.V, c->d_persist[k], P);
    } else if (lb >= c->min_lvl) {
        launch_bottom(c->dim, c->stream, m.T, m.V, c->d_bottom[k], bottom_args(m, lb), c->bottom_lds[k],
                      c->bottom_threads);
    } else {
        gsrb_level(m, c->min_lvl, m.o->n_cycle_base);
    }
    for (int lvl = low + 1; lvl <= highest_lvl; lvl++) correct_relax(m, lvl, m.o->n_cycle_up, in_copy[lvl - c->min_lvl]);
    if (set_residual) {
        if (lb >= c->min_lvl && c->res_upper > c->highest_lvl) {
            if (persist && !res_inside) {
                int64_t most = 0;
                for (int li = 0; li <= lb - c->min_lvl; li++) {
                    const LevelArgs &L = c->h_persist[k].lv[li];
                    most = std::max<int64_t>(most, (int64_t)L.n * (c->dim == 2 ? L.nc * L.nc : L.nc * L.nc * L.nc));
                }
                launch_residual_levels(c->dim, c->stream, m.T, m.V, c->d_persist[k], lb - c->min_lvl, most);
            } else if (!persist) {
                launch_bottom_residual(c->dim, c->stream, m.T, m.V, c->d_bottom[k], bottom_args(m, lb),
                                       c->bottom_threads);
            }
            for (int lvl = lb + 1; lvl <= highest_lvl; lvl++) residual_level(m, lvl);
        } else {
            for (int lvl = c->min_lvl; lvl <= highest_lvl; lvl++) residual_level(m, lvl);
        }
    }
}

// Run a V-cycle, from a cached HIP graph when enabled.  In timing mode the
// graph carries event records around its colour passes; the stream is then
// synchronised after each launch to read them before the next replay.
// `bound` and `res_upper` select the trees of a forest: the V-cycle runs on
// trees whose highest level is >= bound and stores the residual of those
// below res_upper (one tree: bound = min_lvl, res_upper above the highest).
template <typename R>
int run_vcycle(MgT<R> &m, int highest_lvl, int set_residual, int bound, int res_upper)
{
    afg_ctx *c = m.c;
    struct Scope {   // the selection holds for this V-cycle only
        afg_ctx *c;
        ~Scope()
        {
            c->bound = c->min_lvl;
            c->res_upper = c->highest_lvl + 1;
        }
    } scope{c};
    c->bound = bound;
    c->res_upper = std::min(res_upper, c->highest_lvl + 1);
    if (c->persist_wgs > 0) c->persist_ran = true;
    if (!c->use_graphs) {
        vcycle_body(m, highest_lvl, set_residual);
        return 0;
    }
    const afg_mg_opts *o = m.o;
    // In timing mode each key has TIMING_RING graph copies, each with its own
    // events, used in turn: a copy's events are read before it is replayed
    // again, which by then has long finished, so the stream never stalls.
    constexpr unsigned TIMING_RING = 8;
    auto key = std::make_tuple(highest_lvl, set_residual, o->i_phi, o->i_rhs, o->i_tmp,
                               o->i_eps + 1000 * o->i_lsf + 1000000 * o->i_bval,
                               o->op + 10 * (int)sizeof(R) + 1000 * o->rstr, o->n_cycle_down * 1000000 + o->n_cycle_up * 1000,
                               o->n_cycle_base,
                               c->timing ? 1 + (int)(c->timing_ring++ % TIMING_RING) : 0,
                               c->bound * 1000 + c->res_upper);
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        hipGraph_t g = nullptr;
        c->cap_ev.clear();
        c->cap_blocks.clear();
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        c->capturing = true;
        vcycle_body(m, highest_lvl, set_residual);
        c->capturing = false;
        HIP_TRY(c, hipStreamEndCapture(c->stream, &g));
        GraphRec rec;
        hipError_t e = hipGraphInstantiate(&rec.exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        rec.ev.swap(c->cap_ev);
        rec.blocks.swap(c->cap_blocks);
        if (e != hipSuccess) {
            destroy_graph(rec);
            return c->fail(AFG_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
        it = c->graphs.emplace(key, std::move(rec)).first;
    }
    if (int e = harvest_graph(c, it->second)) return e;
    HIP_TRY(c, hipGraphLaunch(it->second.exec, c->stream));
    it->second.pending = !it->second.ev.empty();
    return 0;
}

// mg$D_fas_fmg (m_aX_mg.f90:164-203)
template <typename R>
int fas_fmg(MgT<R> &m, int set_residual, int have_guess)
{
    afg_ctx *c = m.c;
    if (have_guess) {
        for (int lvl = c->highest_lvl; lvl >= c->min_lvl + 1; lvl--) set_coarse_phi_rhs(m, lvl);
    } else {
        init_phi_rhs(m);
    }
    for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++) {
        const LevelArgs L = level_args(c, lvl);
        launch_copy_boxes(c->dim, c->stream, m.T, L.slot0, L.n, L.nc, m.V.phi, m.V.tmp);
        if (lvl > c->min_lvl) {
            // correct_children + fill_gc_phi (m_aX_mg.f90:193-196)
            correct_relax(m, lvl, 0);
        }
        // the V-cycle of FMG step lvl runs on the trees reaching lvl; a tree
        // stores its residual in its last step, lvl == its highest level
        if (int e = run_vcycle(m, lvl, set_residual ? 1 : 0, lvl, lvl + 1)) return e;
    }
    return 0;
}

int check_level(afg_ctx *c, int lvl, bool need_parent)
{
    if (lvl < c->min_lvl || lvl > c->highest_lvl)
        return c->fail(AFG_ERR_ARG, "level %d outside %d..%d", lvl, c->min_lvl, c->highest_lvl);
    if (need_parent && lvl == c->min_lvl)
        return c->fail(AFG_ERR_ARG, "level %d has no parent level", lvl);
    return 0;
}

// The bottom kernel takes the lowest levels while each holds at most
// bottom_cells cells (per precision) and all of them, with their topology tables, fit in the
// LDS of one workgroup: phi, rhs, tmp (and eps, for the auto operators) of
// those levels are copied in, so every access of the kernel is an LDS access.
// Levels of at most bottom_wave_cells cells each, from the lowest up, run in
// one wavefront.  Four tables are kept: FP64 or FP32 values, 3 or 4
// variables in LDS; each has its own highest level.
int choose_bottom(afg_ctx *c)
{
    data_written(c);   // levels may move between the bottom kernel and the per-level passes
    const int nl = (int)c->levels.size();
    int64_t max_threads = 64;
    for (int k = 0; k < 4; k++) {
        int top = -1;
        // cylindrical trees run every level with per-level launches
        for (int li = 0; li < nl && li < MAX_BOTTOM_LEVELS && !c->cyl; li++) {
            const Level &L = c->levels[li];
            const int64_t cells = (int64_t)L.n * (c->dim == 2 ? L.nc * L.nc : L.nc * L.nc * L.nc);
            if (cells > c->bottom_cells[k >= 2]) break;
            top = li;
        }
        const int nv = 3 + (k & 1);
        const int64_t elem = k >= 2 ? (int64_t)sizeof(float) : (int64_t)sizeof(double);
        // the highest level tk <= top whose tables and data (levels 0..tk) fit
        int tk = -1, n_slots = 0, n_bc = 0, n_par = 0, table_bytes = 0;
        int64_t per_var = 0;   // values of one variable over the levels
        for (int li = 0; li <= top; li++) {
            const Level &L = c->levels[li];
            const int ns = L.slot0 + L.n;
            int nb = 0;
            for (const BcRec &r : c->bc) nb += r.slot < ns;
            const int np = n_par + list_args(c, L.lvl, 0).n;
            const int tb = (int)((((int64_t)ns * (c->nnb + 2) + 2 * nb + np) * 4 + 2 * ns + 7) / 8 * 8);
            const int64_t pv = per_var + (int64_t)L.n * box_len(c->dim, L.nc);
            if (tb + pv * nv * elem > c->lds_limit) break;
            tk = li;
            n_slots = ns;
            n_bc = nb;
            n_par = np;
            table_bytes = tb;
            per_var = pv;
        }
        BottomLevels h;
        std::memset(&h, 0, sizeof h);
        h.n_slots = n_slots;
        h.n_bc = n_bc;
        h.n_par = n_par;
        h.table_bytes = table_bytes;
        h.wave_top = -1;
        int par_off = 0;
        for (int li = 0; li <= tk; li++) {
            const Level &L = c->levels[li];
            h.lv[li] = level_args(c, L.lvl);
            h.par[li] = list_args(c, L.lvl, 0);
            h.par_off[li] = par_off;
            par_off += h.par[li].n;
            h.box_len[li] = box_len(c->dim, L.nc);
            h.in_lds[li] = 1;
            if ((int64_t)L.n * (c->dim == 2 ? L.nc * L.nc : L.nc * L.nc * L.nc) <= c->bottom_wave_cells &&
                h.wave_top == li - 1)
                h.wave_top = li;
            max_threads = std::max<int64_t>(max_threads, (int64_t)L.n * h.box_len[li] / 2);
        }
        for (int v = 0; v < 4; v++) {   // variable-major layout over the levels
            int64_t off = v * per_var;
            for (int li = 0; li <= tk; li++) {
                h.lds_off[v][li] = (int)off;
                off += (int64_t)c->levels[li].n * h.box_len[li];
            }
        }
        c->bottom_lvl[k] = tk >= 0 ? c->levels[tk].lvl : -1000;
        c->bottom_lds[k] = tk >= 0 ? (size_t)table_bytes + (size_t)(per_var * nv * elem) : 0;
        c->h_bottom[k] = h;
    }
    c->bottom_threads = (int)std::min<int64_t>(BOTTOM_BLOCK, (max_threads + 63) / 64 * 64);
    const size_t lds_max = *std::max_element(c->bottom_lds, c->bottom_lds + 4);
    if (lds_max > 0 && bottom_prepare(c->dim, lds_max) != 0)
        return c->fail(AFG_ERR_HIP, "cannot allow %zu bytes of LDS for the bottom kernel", lds_max);
    for (auto &g : c->graphs) destroy_graph(g.second);
    c->graphs.clear();
    return build_base(c);
}

// The cells of the lowest level as the base relaxation reads them
// (BaseCell): colour 1, then colour 2 (n[0], n[1] of them); positions in its
// LDS copy (lds) or in the HBM pools.  fast: every box has 2^dim cells and
// physical boundaries on all faces.
int base_cell_list(afg_ctx *c, bool lds, std::vector<BaseCell> &all, int n[2], bool &fast)
{
    const int dim = c->dim, nnb = c->nnb;
    const Level &L0 = c->levels[0];
    const int nc = L0.nc, s = nc + 2;
    const int vstride = lds ? box_len(dim, nc) : c->stride, vslot0 = lds ? L0.slot0 : 0;
    std::vector<BaseCell> cells[2];
    for (int slot = L0.slot0; slot < L0.slot0 + L0.n; slot++) {
        for (int kk = dim == 3 ? 1 : 0; kk <= (dim == 3 ? nc : 0); kk++)
            for (int j = 1; j <= nc; j++)
                for (int i = 1; i <= nc; i++) {
                    const int x[3] = {i, j, kk};
                    BaseCell d;
                    std::memset(&d, 0, sizeof d);
                    d.slot = slot;
                    d.pos = (slot - vslot0) * vstride + i + s * (j + s * kk);
                    const int dq[6] = {-1, 1, -s, s, -s * s, s * s};
                    for (int m = 0; m < nnb; m++) {
                        const int dn = m >> 1, high = m & 1;
                        if (x[dn] != (high ? nc : 1)) {   // neighbour inside the box
                            d.fly[m] = d.pos + dq[m];
                            continue;
                        }
                        const int nid = c->h_nb[(size_t)slot * nnb + m];
                        if (nid > 0) {   // sides_from_nb: the neighbour's mirrored cell
                            int y[3] = {i, j, kk};
                            y[dn] = high ? 1 : nc;
                            const int ns = c->id2slot[nid];
                            d.fly[m] = (ns - vslot0) * vstride + y[0] + s * (y[1] + s * y[2]);
                        } else if (nid < 0) {   // bc_to_gc with the captured value b
                            const BcRec &r = c->bc[c->bc_of[(size_t)slot * nnb + m]];
                            double b = 0.0;
                            if (r.voff >= 0) {
                                const int t1 = dn == 0 ? j : i, t2 = dn == 2 ? j : kk;
                                b = c->bc_vals[r.voff + (t1 - 1) + (dim == 3 ? nc * (t2 - 1) : 0)];
                            }
                            if (r.type == AFG_BC_DIRICHLET) {
                                d.fly[m] = -1;
                                d.K[m] = 2.0 * b;
                            } else {
                                const double c2 = L0.dr * (high ? 1.0 : -1.0);
                                d.fly[m] = -2;
                                d.K[m] = c2 * b;
                            }
                        } else {
                            return c->fail(AFG_ERR_ARG, "refinement boundary on the lowest level");
                        }
                    }
                    cells[(i + j + kk) & 1 ? 0 : 1].push_back(d);
                }
    }
    all = cells[0];
    all.insert(all.end(), cells[1].begin(), cells[1].end());
    n[0] = (int)cells[0].size();
    n[1] = (int)cells[1].size();
    fast = nc == 2;
    for (int slot = L0.slot0; fast && slot < L0.slot0 + L0.n; slot++)
        for (int m = 0; m < nnb; m++) fast = fast && c->h_nb[(size_t)slot * nnb + m] < 0;
    return 0;
}

// The cell lists of the lowest level for the bottom kernel (BaseCell), for
// both tables, and the tables' upload.  Positions follow where the table
// keeps level 0: its LDS copy or the HBM pool.
int build_base(afg_ctx *c)
{
    const Level &L0 = c->levels[0];
    const int ncells = L0.n * (c->dim == 2 ? L0.nc * L0.nc : L0.nc * L0.nc * L0.nc);
    if (ncells > c->base_cap) {
        for (int k = 0; k < 4; k++) {
            ctx_free(c, c->d_base[k]);
            c->d_base[k] = nullptr;
            HIP_TRY(c, ctx_malloc(c, &c->d_base[k], (size_t)ncells * sizeof(BaseCell)));
        }
        c->base_cap = ncells;
    }
    if (int e_ = stream_sync(c)) return e_;
    for (int k = 0; k < 4; k++) {
        BottomLevels &h = c->h_bottom[k];
        if (c->bottom_lvl[k] < c->min_lvl) {
            h.base = nullptr;
            h.base_n[0] = h.base_n[1] = 0;
            HIP_H2D(c, c->d_bottom[k], &h, sizeof h);
            continue;
        }
        std::vector<BaseCell> all;
        bool fast = false;
        if (int e = base_cell_list(c, h.in_lds[0] != 0, all, h.base_n, fast)) return e;
        HIP_H2D(c, c->d_base[k], all.data(), all.size() * sizeof(BaseCell));
        h.base_fast = fast ? 1 : 0;
        h.base = c->d_base[k];
        HIP_H2D(c, c->d_bottom[k], &h, sizeof h);
    }
    return choose_persist(c);
}

// The persistent kernel's tables: it takes the lowest levels while each holds
// at most persist_cells cells (per precision; no LDS bound, its data stay in
// the HBM pools); the tables list those levels, the parent list of each and
// the lowest level's cells at their positions in the pools (the same for all
// four tables).
int choose_persist(afg_ctx *c)
{
    const int nl = (int)c->levels.size();
    bool any = false;
    for (int k = 0; k < 4; k++) {
        int top = -1;
        int64_t most = 0;
        // cylindrical trees run every level with per-level launches
        for (int li = 0; li < nl && li < MAX_BOTTOM_LEVELS && !c->cyl; li++) {
            const Level &L = c->levels[li];
            const int64_t cells = (int64_t)L.n * (c->dim == 2 ? L.nc * L.nc : L.nc * L.nc * L.nc);
            if (cells > c->persist_cells[k >= 2]) break;
            top = li;
            most = std::max(most, cells);
        }
        BottomLevels h;
        std::memset(&h, 0, sizeof h);
        h.wave_top = -1;
        for (int li = 0; li <= top; li++) {
            const Level &L = c->levels[li];
            h.lv[li] = level_args(c, L.lvl);
            h.par[li] = list_args(c, L.lvl, 0);
            h.box_len[li] = box_len(c->dim, L.nc);
        }
        c->persist_lvl[k] = top >= 0 ? c->levels[top].lvl : -1000;
        c->persist_max_cells[k] = most;
        c->h_persist[k] = h;
        any = any || top >= 0;
    }
    if (any) {
        std::vector<BaseCell> all;
        int n[2] = {0, 0};
        bool fast = false;
        if (int e = base_cell_list(c, false, all, n, fast)) return e;
        if ((int)all.size() > c->pbase_cap) {
            ctx_free(c, c->d_pbase);
            c->d_pbase = nullptr;
            c->pbase_cap = 0;
            HIP_TRY(c, ctx_malloc(c, &c->d_pbase, all.size() * sizeof(BaseCell)));
            c->pbase_cap = (int)all.size();
        }
        HIP_H2D(c, c->d_pbase, all.data(), all.size() * sizeof(BaseCell));
        for (int k = 0; k < 4; k++) {
            BottomLevels &h = c->h_persist[k];
            h.base = c->d_pbase;
            h.base_n[0] = n[0];
            h.base_n[1] = n[1];
            h.base_fast = fast ? 1 : 0;
        }
    }
    for (int k = 0; k < 4; k++) HIP_H2D(c, c->d_persist[k], &c->h_persist[k], sizeof(BottomLevels));
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI

// The autotuner of the tiled colour pass: on every level with more than one
// tile shape, each shape runs `reps` times two passes (two single passes,
// and one fused pair; face ghosts computed) on a copy of phi, three times;
// the shape and mode with the least time are kept for the level (the
// current ones unless another is more than 3% faster).  Every shape and
// mode computes the same values, so the choice changes speed only.
template <typename R>
static int tune_tiles(afg_ctx *c, const afg_mg_opts *o, int reps, double *ms, const double *phi64)
{
    MgT<R> m;
    if (int e = make_mg(c, o, &m)) return e;
    const size_t n = (size_t)c->n_slots * c->stride;
    R *scratch = nullptr;
    HIP_TRY(c, ctx_malloc(c, &scratch, n * sizeof(R)));
    MgVarsT<R> V = m.V;
    V.phi = scratch;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int err = 0;
    const int bound = c->bound;
    c->bound = c->min_lvl;      // every tree's boxes
    auto run = [&]() -> int {
        if constexpr (sizeof(R) == 4)
            launch_to_fp32(c->stream, (long long)n, phi64, scratch);
        else
            HIP_TRY(c, hipMemcpyAsync(scratch, phi64, n * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipEventCreate(&e0));
        HIP_TRY(c, hipEventCreate(&e1));
        for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++) {
            Level &Lv = c->levels[lvl - c->min_lvl];
            const int nv = tile_shapes(Lv.nc);
            LevelArgs L = level_args(c, lvl);
            if (nv < 2 || L.n == 0) continue;
            const int p = sizeof(R) == 4;
            const int dflt = Lv.tile[p] + (Lv.pairs[p] ? 3 : 0);
            // time of two passes, timed as blocks of four: single passes
            // (mode 0; on the colour-split copy of phi a first, two middle
            // and a last pass, where the relaxation blocks use it) or two
            // fused pairs (mode 1: the copy of phi to the second pool and
            // back, the other levels from the tree's phi), per shape
            double per[6];
            for (int mode = 0; mode < 2; mode++)
                for (int v = 0; v < nv; v++) {
                    L.tile = v;
                    auto step = [&](int) {
                        if (mode) {
                            launch_gsrb_pair(c->dim, c->stream, m.T, L, m.V, scratch, m.phi2, 1, 1);
                            launch_gsrb_pair(c->dim, c->stream, m.T, L, m.V, m.phi2, scratch, 3, 1);
                        } else if (launch_gsrb_split(c->dim, c->stream, m.T, L, V, 1, 0, 1)) {
                            launch_gsrb_split(c->dim, c->stream, m.T, L, V, 2, 1, 1);
                            launch_gsrb_split(c->dim, c->stream, m.T, L, V, 3, 1, 1);
                            launch_gsrb_split(c->dim, c->stream, m.T, L, V, 4, 2, 1);
                        } else {
                            for (int rb = 1; rb <= 4; rb++) launch_gsrb_pass(c->dim, c->stream, m.T, L, V, rb, 1);
                        }
                    };
                    for (int r = 0; r < 2; r++) step(r);
                    float t_min = 0;
                    for (int trial = 0; trial < 3; trial++) {
                        HIP_TRY(c, hipEventRecord(e0, c->stream));
                        for (int r = 0; r < reps; r++) step(r);
                        HIP_TRY(c, hipEventRecord(e1, c->stream));
                        HIP_TRY(c, hipEventSynchronize(e1));
                        float t = 0;
                        HIP_TRY(c, hipEventElapsedTime(&t, e0, e1));
                        if (trial == 0 || t < t_min) t_min = t;
                    }
                    per[3 * mode + v] = (double)t_min / reps / 2;
                    if (ms) ms[(lvl - c->min_lvl) * 6 + 3 * mode + v] = per[3 * mode + v];
                }
            // the fastest (mode, shape), unless the current one is within 3%
            // of it (timing noise would otherwise flip choices between runs)
            int best = 0;
            for (int k = 1; k < 6; k++)
                if (k % 3 < nv && per[k] < per[best]) best = k;
            const int pick = per[dflt] <= 1.03 * per[best] ? dflt : best;
            Lv.tile[p] = pick % 3;
            Lv.pairs[p] = pick >= 3;
        }
        return 0;
    };
    err = run();
    c->bound = bound;
    hipStreamSynchronize(c->stream);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    ctx_free(c, scratch);
    for (auto &g : c->graphs) destroy_graph(g.second);
    c->graphs.clear();
    return err;
}

// The autotuner of the bottom kernel's cell limit (per precision): a few
// V-cycles with each candidate limit, timed with HIP events; phi, rhs and tmp
// are saved before and restored after, so the tree's data is unchanged.
template <typename R>
static int tune_bottom(afg_ctx *c, const afg_mg_opts *o, int reps, double *ms)
{
    MgT<R> m;
    if (int e = make_mg(c, o, &m)) return e;
    const size_t n = (size_t)c->n_slots * c->stride;
    R *saved = nullptr;
    HIP_TRY(c, ctx_malloc(c, &saved, 3 * n * sizeof(R)));
    R *vars[3] = {m.V.phi, m.V.rhs, m.V.tmp};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int p = sizeof(R) == 4;
    const int64_t keep = c->bottom_cells[p];
    const int64_t cand[4] = {512, 2048, 8192, 32768};
    double t_best = 0;
    int64_t best = keep;
    auto run = [&]() -> int {
        for (int v = 0; v < 3; v++)
            HIP_TRY(c, hipMemcpyAsync(saved + v * n, vars[v], n * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipEventCreate(&e0));
        HIP_TRY(c, hipEventCreate(&e1));
        for (int i = 0; i < 4; i++) {
            c->bottom_cells[p] = cand[i];
            if (int e = choose_bottom(c)) return e;
            for (int r = 0; r < 2; r++)
                if (int e = run_vcycle(m, c->highest_lvl, 0, c->min_lvl, c->highest_lvl + 1)) return e;
            HIP_TRY(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; r++)
                if (int e = run_vcycle(m, c->highest_lvl, 0, c->min_lvl, c->highest_lvl + 1)) return e;
            HIP_TRY(c, hipEventRecord(e1, c->stream));
            HIP_TRY(c, hipEventSynchronize(e1));
            float t = 0;
            HIP_TRY(c, hipEventElapsedTime(&t, e0, e1));
            if (ms) ms[i] = (double)t / reps;
            if (i == 0 || t < t_best) {
                t_best = t;
                best = cand[i];
            }
        }
        for (int v = 0; v < 3; v++)
            HIP_TRY(c, hipMemcpyAsync(vars[v], saved + v * n, n * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
        return 0;
    };
    const int err = run();
    c->bottom_cells[p] = err ? keep : best;
    const int e2 = choose_bottom(c);
    hipStreamSynchronize(c->stream);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    ctx_free(c, saved);
    return err ? err : e2;
}

// A box as the tree update matches it between the old and the new tree: its
// tree, level and index (ids may differ: a$D_tidy_up renumbers them,
// m_aX_core.f90:238-314, and freed ids are handed out again).
struct BoxKey {
    int tree, lvl, ix[3];
    bool operator==(const BoxKey &o) const
    {
        return tree == o.tree && lvl == o.lvl && ix[0] == o.ix[0] && ix[1] == o.ix[1] && ix[2] == o.ix[2];
    }
};

struct BoxKeyHash {
    size_t operator()(const BoxKey &k) const
    {
        uint64_t h = 0xcbf29ce484222325ull;
        for (int v : {k.tree, k.lvl, k.ix[0], k.ix[1], k.ix[2]}) h = (h ^ (uint32_t)v) * 0x100000001b3ull;
        return (size_t)(h ^ (h >> 31));
    }
};

static BoxKey box_key(const afg_ctx *c, int s)
{
    const int32_t *ix = &c->slot_ix[(size_t)s * 3];
    return BoxKey{c->slot_tree[s], c->slot_lvl[s], {ix[0], ix[1], ix[2]}};
}

namespace afg {
// the message of a failed call that has no context (afg_last_error(NULL))
void set_call_error(const char *msg) { g_create_error = msg; }
}  // namespace afg

extern "C" {

int afg_ctx_create(int dim, int n_cell, int n_var, afg_ctx **out)
{
    if (!out) return AFG_ERR_ARG;
    *out = nullptr;
    if (dim != 2 && dim != 3) {
        g_create_error = "dim must be 2 or 3";
        return AFG_ERR_ARG;
    }
    // a$D_init checks (m_aX_core.f90:61-63)
    if (n_cell < 2 || (n_cell & 1) || n_cell > 1000) {
        g_create_error = "n_cell should be even, >= 2 and <= 1000";
        return AFG_ERR_ARG;
    }
    if (n_var <= 0) {
        g_create_error = "n_var_cell should be > 0";
        return AFG_ERR_ARG;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        g_create_error = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return AFG_ERR_HIP;
    }
    afg_ctx *c = new afg_ctx;
    {
        int a = 0, b = 0;
        hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, dev);
        c->lds_limit = std::max(a, b) > 0 ? std::max(a, b) : 65536;
    }
    c->dim = dim;
    c->nc = n_cell;
    c->n_var = n_var;
    c->nch = 1 << dim;
    c->nnb = 2 * dim;
    c->device = dev;
    // each box array padded to a multiple of 16 doubles (128 B)
    c->stride = (box_len(dim, n_cell) + 15) / 16 * 16;
    c->fstride = (dim * (n_cell + 1) * (dim == 3 ? n_cell * n_cell : n_cell) + 15) / 16 * 16;
    c->d_fc.assign(AFG_MAX_FC, nullptr);
    c->d_var.assign(n_var, nullptr);
    c->d_var32.assign(n_var, nullptr);
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // the allocation-free entry point's scratch array on this device (the
    // current one): its address, looked up once per device
    if (e == hipSuccess) e = scratch_address(dev, &c->scratch);
    if (e == hipSuccess) e = ctx_malloc(c, &c->d_out, 8 * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_out, 8 * sizeof(double), hipHostMallocDefault);
    // the persistent kernel's barrier counters start at zero (it leaves them so)
    if (e == hipSuccess) e = ctx_malloc(c, &c->d_psync, (4 + PERSIST_MAX_RANKS) * sizeof(unsigned));
    // (on the context's own stream: a NULL-stream call here would disturb a
    // graph capture that another thread's context may be running)
    if (e == hipSuccess) e = hipMemsetAsync(c->d_psync, 0, (4 + PERSIST_MAX_RANKS) * sizeof(unsigned), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_psync, 4 * sizeof(unsigned), hipHostMallocDefault);
    {
        // contexts take the XCDs in turn, so that the persistent kernels of
        // contexts that run at once (on their own streams) use different L2s
        static std::atomic<unsigned> n_created{0};
        c->persist_xcd = (int)(n_created++ % 8);
    }
    if (e != hipSuccess) {
        g_create_error = std::string("HIP set-up: ") + hipGetErrorString(e);
        afg_ctx_destroy(c);
        return AFG_ERR_HIP;
    }
    *out = c;
    return 0;
}

int afg_ctx_destroy(afg_ctx *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    free_tree(c);
    for (auto &ev : c->ev_used) {
        hipEventDestroy(ev.first);
        hipEventDestroy(ev.second);
    }
    for (auto &ev : c->ev_free) {
        hipEventDestroy(ev.first);
        hipEventDestroy(ev.second);
    }
    ctx_free(c, c->d_out);
    ctx_free(c, c->d_psync);
    ctx_free(c, c->d_ptrace);
    if (c->h_out) hipHostFree(c->h_out);
    if (c->h_psync) hipHostFree(c->h_psync);
    if (c->h_stage) hipHostFree(c->h_stage);
    for (hipEvent_t ev : c->xfer_ev)
        if (ev) hipEventDestroy(ev);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

const char *afg_last_error(const afg_ctx *c)
{
    return c ? c->err.c_str() : g_create_error.c_str();
}

int afg_sync(afg_ctx *c)
{
    if (!c) return AFG_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    if (int e_ = stream_sync(c)) return e_;
    return 0;
}

int afg_set_tree(afg_ctx *c, int n_boxes, const int32_t *recs, double dr_base)
{
    return afg_set_forest(c, n_boxes, recs, dr_base, nullptr);
}

static int build_cyl(afg_ctx *c);

int afg_set_forest(afg_ctx *c, int n_boxes, const int32_t *recs, double dr_base, const int32_t *tree_of)
{
    if (!c) return AFG_ERR_ARG;
    if (n_boxes < 1 || !recs) return c->fail(AFG_ERR_ARG, "empty tree");
    if (!(dr_base > 0)) return c->fail(AFG_ERR_ARG, "dr_base must be > 0");
    HIP_TRY(c, hipSetDevice(c->device));
    if (int e_ = stream_sync(c)) return e_;
    free_tree(c);
    c->mixed_ready = c->mixed_fmg = false;
    c->tags_set = false;
    const int dim = c->dim, nch = c->nch, nnb = c->nnb;
    const int rl = 3 + dim + nch + nnb;
    int max_id = 0, min_lvl = 1 << 30, hi_lvl = -(1 << 30);
    for (int r = 0; r < n_boxes; r++) {
        const int32_t *q = recs + (size_t)r * rl;
        if (q[0] < 1) return c->fail(AFG_ERR_ARG, "record %d: box id %d < 1", r, q[0]);
        max_id = std::max(max_id, q[0]);
        min_lvl = std::min(min_lvl, q[1]);
        hi_lvl = std::max(hi_lvl, q[1]);
    }
    if (hi_lvl - min_lvl > 60) return c->fail(AFG_ERR_ARG, "level range %d..%d too large", min_lvl, hi_lvl);
    std::vector<int32_t> rec_of(max_id + 1, -1);
    for (int r = 0; r < n_boxes; r++) {
        const int id = recs[(size_t)r * rl];
        if (rec_of[id] >= 0) return c->fail(AFG_ERR_ARG, "box id %d appears twice", id);
        rec_of[id] = r;
    }
    // trees of a forest: each tree's highest level
    int n_trees = 1;
    if (tree_of) {
        for (int r = 0; r < n_boxes; r++) {
            if (tree_of[r] < 0 || tree_of[r] >= n_boxes)
                return c->fail(AFG_ERR_ARG, "record %d: tree index %d out of range", r, tree_of[r]);
            n_trees = std::max(n_trees, tree_of[r] + 1);
        }
    }
    auto tree = [&](int r) { return tree_of ? tree_of[r] : 0; };
    std::vector<int> tree_hi(n_trees, -(1 << 30));
    for (int r = 0; r < n_boxes; r++) tree_hi[tree(r)] = std::max(tree_hi[tree(r)], recs[(size_t)r * rl + 1]);
    for (int t = 0; t < n_trees; t++)
        if (tree_hi[t] < min_lvl) return c->fail(AFG_ERR_ARG, "tree %d has no boxes", t);
    // group by level, then by tree (deepest trees first, so that the trees
    // reaching a level are a prefix of every level), Morton order within a
    // tree's level (the key of a$D_tidy_up, m_aX_core.f90:272-279:
    // morton_from_ix$D(ix - 1))
    const int nl = hi_lvl - min_lvl + 1;
    struct Key {
        int neg_hi, tree;
        uint64_t morton;
        int rec;
    };
    std::vector<std::vector<Key>> per(nl);
    for (int r = 0; r < n_boxes; r++) {
        const int32_t *q = recs + (size_t)r * rl;
        uint64_t key;
        if (dim == 2)
            key = spread2(q[2] - 1) | (spread2(q[3] - 1) << 1);
        else
            key = spread3(q[2] - 1) | (spread3(q[3] - 1) << 1) | (spread3(q[4] - 1) << 2);
        per[q[1] - min_lvl].push_back({-tree_hi[tree(r)], tree(r), key, r});
    }
    c->levels.assign(nl, Level());
    c->id2slot.assign(max_id + 1, -1);
    c->slot2id.assign(n_boxes, 0);
    c->slot_lvl.assign(n_boxes, 0);
    c->slot_nc.assign(n_boxes, 0);
    c->slot_ix.assign((size_t)n_boxes * 3, 0);
    int slot = 0;
    for (int li = 0; li < nl; li++) {
        const int lvl = min_lvl + li;
        if (per[li].empty()) return c->fail(AFG_ERR_ARG, "level %d has no boxes", lvl);
        std::stable_sort(per[li].begin(), per[li].end(), [](const Key &a, const Key &b) {
            if (a.neg_hi != b.neg_hi) return a.neg_hi < b.neg_hi;
            if (a.tree != b.tree) return a.tree < b.tree;
            return a.morton < b.morton;
        });
        Level &L = c->levels[li];
        L.lvl = lvl;
        // a$D_n_cell (m_aX_utils.f90:676-686)
        L.nc = lvl >= 1 ? c->nc : c->nc >> (1 - lvl);
        if (L.nc < 1 || (lvl < 1 && (c->nc % (1 << (1 - lvl))) != 0))
            return c->fail(AFG_ERR_ARG, "level %d: n_cell %d cannot be coarsened to it", lvl, c->nc);
        if (L.nc < 2 || (L.nc & 1))
            return c->fail(AFG_ERR_UNSUPPORTED, "level %d: n_cell %d not supported (coarse levels need an even n_cell >= 2)", lvl, L.nc);
        // default tile shapes (afg_autotune refines them): 3D boxes of 16^3
        // cells take 1024 threads, two cells of the colour per thread
        L.tile[0] = L.tile[1] = dim == 3 && L.nc == 16 ? 2 : 0;
        // fused pass pairs by default in 2D (faster there; in 3D the
        // neighbour work they repeat costs more than the launch they save)
        L.pairs[0] = L.pairs[1] = dim == 2;
        // box%dr = dr_base * 0.5**(lvl-1) (m_aX_core.f90:196, halved per child :764)
        L.dr = dr_base * std::pow(0.5, lvl - 1);
        L.dx2 = L.dr * L.dr;
        L.inv_dr_sq = 1 / (L.dr * L.dr);
        L.slot0 = slot;
        L.n = (int)per[li].size();
        for (auto &kr : per[li]) {
            const int id = recs[(size_t)kr.rec * rl];
            c->id2slot[id] = slot;
            c->slot2id[slot] = id;
            c->slot_lvl[slot] = lvl;
            c->slot_nc[slot] = L.nc;
            for (int d = 0; d < dim; d++) c->slot_ix[(size_t)slot * 3 + d] = recs[(size_t)kr.rec * rl + 2 + d];
            slot++;
        }
    }
    c->n_slots = slot;
    c->min_lvl = min_lvl;
    c->highest_lvl = hi_lvl;
    c->n_trees = n_trees;
    c->tree_hi = tree_hi;
    c->min_tree_hi = *std::min_element(tree_hi.begin(), tree_hi.end());
    c->bound = min_lvl;
    c->res_upper = hi_lvl + 1;
    std::vector<int> slot_tree(slot);
    for (int r = 0; r < n_boxes; r++) slot_tree[c->id2slot[recs[(size_t)r * rl]]] = tree(r);
    c->slot_tree = slot_tree;
    c->max_id = max_id;
    c->dr_base = dr_base;

    // per-slot tables
    std::vector<int32_t> nbv((size_t)slot * nnb), par(slot, -1), choff(slot, 0);
    c->h_nb.assign((size_t)slot * nnb, 0);
    c->bc.clear();
    c->bc_vals.clear();
    c->bc_of.assign((size_t)slot * nnb, -1);
    std::vector<int32_t> lists;
    for (int s = 0; s < slot; s++) {
        const int32_t *q = recs + (size_t)rec_of[c->slot2id[s]] * rl;
        const int lvl = q[1];
        const int pid = q[2 + dim];
        if (lvl > min_lvl) {
            if (pid < 1 || pid > max_id || c->id2slot[pid] < 0)
                return c->fail(AFG_ERR_ARG, "box %d: parent %d missing", q[0], pid);
            const int ps = c->id2slot[pid];
            if (c->slot_lvl[ps] != lvl - 1) return c->fail(AFG_ERR_ARG, "box %d: parent %d not on level %d", q[0], pid, lvl - 1);
            if (slot_tree[ps] != slot_tree[s]) return c->fail(AFG_ERR_ARG, "box %d: parent %d in another tree", q[0], pid);
            par[s] = ps;
            // a$D_get_child_offset (m_aX_t.f90:362-374)
            int co = 0;
            if (lvl > 1)
                for (int d = 0; d < dim; d++) co |= (((q[2 + d] - 1) & 1) * (c->nc >> 1)) << (10 * d);
            choff[s] = co;
        }
        for (int f = 0; f < nnb; f++) {
            const int nid = q[3 + dim + nch + f];
            c->h_nb[(size_t)s * nnb + f] = nid;
            if (nid > 0) {
                if (nid > max_id || c->id2slot[nid] < 0) return c->fail(AFG_ERR_ARG, "box %d: neighbour %d missing", q[0], nid);
                const int ns = c->id2slot[nid];
                if (c->slot_lvl[ns] != lvl) return c->fail(AFG_ERR_ARG, "box %d: neighbour %d on another level", q[0], nid);
                if (slot_tree[ns] != slot_tree[s]) return c->fail(AFG_ERR_ARG, "box %d: neighbour %d in another tree", q[0], nid);
                nbv[(size_t)s * nnb + f] = ns;
            } else if (nid == 0) {
                if (lvl == min_lvl) return c->fail(AFG_ERR_ARG, "box %d: refinement boundary on the lowest level", q[0]);
                nbv[(size_t)s * nnb + f] = NB_REFINE;
            } else {
                const int rec = (int)c->bc.size();
                c->bc.push_back({s, f, AFG_BC_DIRICHLET, -1});   // a$D_bc_dirichlet_zero
                c->bc_of[(size_t)s * nnb + f] = rec;
                nbv[(size_t)s * nnb + f] = NB_BC0 - rec;
            }
        }
    }
    // a$D_gc_interp reads the parent's neighbour across a refinement boundary
    c->rb_parent_nb = true;
    for (int s = 0; s < slot; s++)
        for (int f = 0; f < nnb; f++)
            if (nbv[(size_t)s * nnb + f] == NB_REFINE && (par[s] < 0 || nbv[(size_t)par[s] * nnb + f] < 0))
                c->rb_parent_nb = false;
    for (int li = 0; li < nl; li++) {
        Level &L = c->levels[li];
        L.parents_off = (int)lists.size();
        for (int s = L.slot0; s < L.slot0 + L.n; s++) {
            const int32_t *q = recs + (size_t)rec_of[c->slot2id[s]] * rl;
            if (q[3 + dim] != 0) lists.push_back(s);   // a$D_has_children
        }
        L.n_parents = (int)lists.size() - L.parents_off;
        L.leaves_off = (int)lists.size();
        for (int s = L.slot0; s < L.slot0 + L.n; s++) {
            const int32_t *q = recs + (size_t)rec_of[c->slot2id[s]] * rl;
            if (q[3 + dim] == 0) lists.push_back(s);
        }
        L.n_leaves = (int)lists.size() - L.leaves_off;
        L.all_off = (int)lists.size();
        for (int s = L.slot0; s < L.slot0 + L.n; s++) lists.push_back(s);
    }
    // boxes, parents and leaves of each level in trees reaching bound b
    for (int li = 0; li < nl; li++) {
        Level &L = c->levels[li];
        L.n_act.assign(nl + 1, 0);
        L.n_par_act.assign(nl + 1, 0);
        L.n_leaf_act.assign(nl + 1, 0);
        for (int b = 0; b <= nl; b++) {
            const int lb = min_lvl + b;
            for (int s2 = L.slot0; s2 < L.slot0 + L.n; s2++) L.n_act[b] += tree_hi[slot_tree[s2]] >= lb;
            for (int k = 0; k < L.n_parents; k++)
                L.n_par_act[b] += tree_hi[slot_tree[lists[L.parents_off + k]]] >= lb;
            for (int k = 0; k < L.n_leaves; k++)
                L.n_leaf_act[b] += tree_hi[slot_tree[lists[L.leaves_off + k]]] >= lb;
        }
    }
    // every box of a level above the lowest must be the child of a parent of
    // the level below (all children of a refined box exist)
    for (int li = 1; li < nl; li++) {
        long long expect = 0;
        const Level &P = c->levels[li - 1];
        for (int k = 0; k < P.n_parents; k++) {
            const int ps = lists[P.parents_off + k];
            const int32_t *q = recs + (size_t)rec_of[c->slot2id[ps]] * rl;
            for (int ch = 0; ch < nch; ch++) {
                const int cid = q[3 + dim + ch];
                if (cid == 0) continue;
                if (cid > max_id || c->id2slot[cid] < 0 || par[c->id2slot[cid]] != ps)
                    return c->fail(AFG_ERR_ARG, "box %d: child %d inconsistent", q[0], cid);
                expect++;
            }
        }
        if (expect != c->levels[li].n)
            return c->fail(AFG_ERR_ARG, "level %d: %d boxes but %lld children of level %d", P.lvl + 1, c->levels[li].n, expect, P.lvl);
    }

    const size_t sn = (size_t)slot;
    HIP_TRY(c, ctx_malloc(c, &c->d_nb, sn * nnb * sizeof(int32_t)));
    HIP_TRY(c, ctx_malloc(c, &c->d_parent, sn * sizeof(int32_t)));
    HIP_TRY(c, ctx_malloc(c, &c->d_choff, sn * sizeof(int32_t)));
    HIP_TRY(c, ctx_malloc(c, &c->d_lists, std::max<size_t>(lists.size(), 1) * sizeof(int32_t)));
    HIP_TRY(c, ctx_malloc(c, &c->d_tag, sn * sizeof(int8_t)));
    HIP_H2D(c, c->d_nb, nbv.data(), sn * nnb * sizeof(int32_t));
    HIP_H2D(c, c->d_parent, par.data(), sn * sizeof(int32_t));
    c->slot_par = par;
    HIP_H2D(c, c->d_choff, choff.data(), sn * sizeof(int32_t));
    HIP_H2D(c, c->d_lists, lists.data(), lists.size() * sizeof(int32_t));
    HIP_TRY(c, hipMemsetAsync(c->d_tag, 0, sn * sizeof(int8_t), c->stream));
    if (int e_ = stream_sync(c)) return e_;
    int maxp = 0;
    for (int li = 0; li < nl; li++) {
        const Level &L = c->levels[li];
        ListArgs a;
        a.n = L.n;
        a.nc = L.nc;
        maxp += extreme_blocks(dim, a);
    }
    c->partial_cap = std::max(maxp, 1);
    {
        std::vector<int8_t> isp(sn, 0);
        for (int s2 = 0; s2 < slot; s2++) {
            const int32_t *q = recs + (size_t)rec_of[c->slot2id[s2]] * rl;
            isp[s2] = q[3 + dim] != 0;   // a$D_has_children
        }
        HIP_TRY(c, ctx_malloc(c, &c->d_is_parent, sn));
        HIP_H2D(c, c->d_is_parent, isp.data(), sn);
    }
    {   // leaf slots grouped by tree, for the per-tree norms
        std::vector<int32_t> tl, toff(n_trees + 1, 0);
        std::vector<std::vector<int32_t>> by(n_trees);
        for (int li = 0; li < nl; li++) {
            const Level &L = c->levels[li];
            for (int k = 0; k < L.n_leaves; k++) {
                const int s2 = lists[L.leaves_off + k];
                if (c->slot_nc[s2] != c->nc)
                    return c->fail(AFG_ERR_ARG, "leaf box %d below level 1", c->slot2id[s2]);
                by[slot_tree[s2]].push_back(s2);
            }
        }
        const int per_box = dim == 2 ? c->nc * c->nc : c->nc * c->nc * c->nc;
        const int per_chunk = std::max(1, 4096 / per_box);
        std::vector<int32_t> chunk;
        for (int t = 0; t < n_trees; t++) {
            toff[t] = (int)chunk.size();
            for (size_t k = 0; k < by[t].size(); k += per_chunk) chunk.push_back((int)(tl.size() + k));
            tl.insert(tl.end(), by[t].begin(), by[t].end());
        }
        toff[n_trees] = (int)chunk.size();
        c->n_chunks = (int)chunk.size();
        chunk.push_back((int)tl.size());
        HIP_TRY(c, ctx_malloc(c, &c->d_tree_leaves, std::max<size_t>(tl.size(), 1) * sizeof(int32_t)));
        HIP_TRY(c, ctx_malloc(c, &c->d_chunk, chunk.size() * sizeof(int32_t)));
        HIP_TRY(c, ctx_malloc(c, &c->d_tree_chunk, toff.size() * sizeof(int32_t)));
        HIP_TRY(c, ctx_malloc(c, &c->d_chunk_max, std::max(c->n_chunks, 1) * sizeof(double)));
        HIP_TRY(c, ctx_malloc(c, &c->d_tree_out, n_trees * sizeof(double)));
        HIP_TRY(c, hipHostMalloc((void **)&c->h_tree_out, n_trees * sizeof(double), hipHostMallocDefault));
        if (!tl.empty())
            HIP_H2D(c, c->d_tree_leaves, tl.data(), tl.size() * sizeof(int32_t));
        HIP_H2D(c, c->d_chunk, chunk.data(), chunk.size() * sizeof(int32_t));
        HIP_H2D(c, c->d_tree_chunk, toff.data(), toff.size() * sizeof(int32_t));
    }
    c->have_tree = true;
    for (int k = 0; k < 4; k++) HIP_TRY(c, ctx_malloc(c, &c->d_bottom[k], sizeof(BottomLevels)));
    for (int k = 0; k < 4; k++) HIP_TRY(c, ctx_malloc(c, &c->d_persist[k], sizeof(BottomLevels)));
    if (int e = choose_bottom(c)) return e;
    HIP_TRY(c, ctx_malloc(c, &c->d_partial, c->partial_cap * sizeof(double)));
    c->bc_dirty = true;
    if (int e = upload_bc(c)) return e;
    if (int e = build_cyl(c)) return e;
    return 0;
}

int afg_set_coord(afg_ctx *c, int coord, double r_base)
{
    if (!c) return AFG_ERR_ARG;
    if (coord != AFG_COORD_XYZ && coord != AFG_COORD_CYL) return c->fail(AFG_ERR_ARG, "unknown coordinate system %d", coord);
    if (coord == AFG_COORD_CYL && c->dim != 2) return c->fail(AFG_ERR_ARG, "a$D_init: cannot have 3d cyl coords");
    if (c->capturing) return c->fail(AFG_ERR_STATE, "coordinate change inside a graph capture");
    HIP_TRY(c, hipSetDevice(c->device));
    if (int e_ = stream_sync(c)) return e_;
    c->cyl = coord == AFG_COORD_CYL;
    c->r_base = r_base;
    for (auto &g : c->graphs) destroy_graph(g.second);
    c->graphs.clear();
    if (!c->have_tree) return 0;
    if (int e = choose_bottom(c)) return e;
    return build_cyl(c);
}

// The tree update behind afg_update_tree and afg_update_forest.  A box of the
// new tree keeps its data when the old tree has a box of the same level and
// ix in the same tree: tree t of the new forest is tree old_tree[t] of the
// old one (identity when old_tree is null; -1: a tree without a predecessor,
// whose boxes start from zeros).
static int update_tree(afg_ctx *c, int n_boxes, const int32_t *recs, double dr_base, const int32_t *tree_of,
                       const int32_t *old_tree, int prolong_order, int n_vars, const int32_t *vars)
{
    if (int e = check_ctx(c)) return e;
    if (c->capturing) return c->fail(AFG_ERR_STATE, "tree update inside a graph capture");
    if (prolong_order < -1 || prolong_order > 2)
        return c->fail(AFG_ERR_ARG, "prolong_order %d not in -1..2", prolong_order);
    if (n_vars < 0 || (n_vars > 0 && !vars)) return c->fail(AFG_ERR_ARG, "invalid variable list");
    for (int k = 0; k < n_vars; k++)
        if (int e = check_var(c, vars[k])) return e;
    if (c->stride & 1) return c->fail(AFG_ERR_UNSUPPORTED, "odd box length %d", c->stride);
    if (int e_ = stream_sync(c)) return e_;

    // the old tree: its boxes, boundary conditions, level settings and pools
    const int nnb = c->nnb, dim = c->dim;
    std::unordered_map<BoxKey, int, BoxKeyHash> old_slot;
    old_slot.reserve((size_t)c->n_slots * 2);
    for (int s = 0; s < c->n_slots; s++) old_slot.emplace(box_key(c, s), s);
    const std::vector<BcRec> old_bc = c->bc;
    const std::vector<int32_t> old_bc_of = c->bc_of;
    const std::vector<double> old_bc_vals = c->bc_vals;
    std::map<int, Level> old_level;
    for (const Level &L : c->levels) old_level[L.lvl] = L;
    std::vector<double *> old_pool = c->d_var, old_fc = c->d_fc;
    std::fill(c->d_var.begin(), c->d_var.end(), nullptr);
    std::fill(c->d_fc.begin(), c->d_fc.end(), nullptr);
    int32_t *d_src = nullptr;
    auto release = [&](int code) {
        for (double *p : old_pool) ctx_free(c, p);
        for (double *p : old_fc) ctx_free(c, p);
        ctx_free(c, d_src);
        return code;
    };
    const int old_n_trees = c->n_trees;
    if (int e = afg_set_forest(c, n_boxes, recs, dr_base, tree_of)) return release(e);
    if (old_tree)
        for (int t = 0; t < c->n_trees; t++)
            if (old_tree[t] < -1 || old_tree[t] >= old_n_trees)
                return release(c->fail(AFG_ERR_ARG, "tree %d: predecessor %d not in -1..%d", t, old_tree[t],
                                       old_n_trees - 1));

    // src[slot]: the box's old slot, -1 for a box the update added; the
    // added boxes level by level (lowest first)
    const int ns = c->n_slots, nl = (int)c->levels.size();
    std::vector<int32_t> src(ns, -1), added;
    std::vector<int> lv_off(nl + 1, 0);
    for (int li = 0; li < nl; li++) {
        const Level &L = c->levels[li];
        lv_off[li] = (int)added.size();
        for (int s = L.slot0; s < L.slot0 + L.n; s++) {
            BoxKey key = box_key(c, s);
            if (old_tree) key.tree = old_tree[key.tree];
            auto it = key.tree < 0 ? old_slot.end() : old_slot.find(key);
            if (it != old_slot.end()) src[s] = it->second;
            else added.push_back(s);
        }
    }
    lv_off[nl] = (int)added.size();
    if (ctx_malloc(c, &d_src, ((size_t)ns + added.size()) * sizeof(int32_t)) != hipSuccess) {
        d_src = nullptr;
        return release(c->fail(AFG_ERR_MEMORY, "tree update: slot map of %d slots", ns));
    }
    int32_t *d_added = d_src + ns;
    if (hipMemcpyAsync(d_src, src.data(), (size_t)ns * sizeof(int32_t), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        (!added.empty() && hipMemcpyAsync(d_added, added.data(), added.size() * sizeof(int32_t),
                                          hipMemcpyHostToDevice, c->stream) != hipSuccess))
        return release(c->fail(AFG_ERR_HIP, "tree update: copying the slot map"));

    // every variable the old tree had: gathered into a new pool (zeros on
    // the added boxes)
    const size_t pool_bytes = (size_t)ns * c->stride * sizeof(double);
    for (int iv = 1; iv <= c->n_var; iv++) {
        if (!old_pool[iv - 1]) continue;
        double *p = nullptr;
        if (ctx_malloc(c, &p, pool_bytes) != hipSuccess)
            return release(c->fail(AFG_ERR_MEMORY, "hipMalloc(%zu) for variable %d", pool_bytes, iv));
        launch_gather(c->stream, c->stride, ns, d_src, old_pool[iv - 1], p);
        c->d_var[iv - 1] = p;
    }
    for (int k = 0; k < AFG_MAX_FC; k++) {
        if (!old_fc[k]) continue;
        double *p = nullptr;
        const size_t fbytes = (size_t)ns * c->fstride * sizeof(double);
        if (ctx_malloc(c, &p, fbytes) != hipSuccess)
            return release(c->fail(AFG_ERR_MEMORY, "hipMalloc(%zu) for face variable %d", fbytes, k + 1));
        launch_gather(c->stream, c->fstride, ns, d_src, old_fc[k], p);
        c->d_fc[k] = p;
    }
    // the listed variables of the added boxes from their parents, level by
    // level (a box's parent may itself be new); added boxes of the lowest
    // level have no parent and keep zeros
    if (prolong_order >= 0) {
        const DevTree T = dev_tree(c);
        for (int k = 0; k < n_vars; k++) {
            double *v = nullptr;
            if (int e = ensure_var(c, vars[k], &v)) return release(e);
            for (int li = 1; li < nl; li++)
                launch_prolong_list(dim, c->stream, T, c->levels[li].nc, c->levels[li - 1].nc,
                                    d_added + lv_off[li], lv_off[li + 1] - lv_off[li], prolong_order, v);
        }
    }
    const hipError_t se = hipStreamSynchronize(c->stream);
    release(0);
    if (se != hipSuccess) return c->fail(AFG_ERR_HIP, "tree update: %s", hipGetErrorString(se));
    if (int e = launched(c, "afg_update_tree")) return e;

    // boundary conditions of the faces that were physical boundaries before
    for (int s = 0; s < ns; s++) {
        if (src[s] < 0) continue;
        for (int f = 0; f < nnb; f++) {
            const int rec = c->bc_of[(size_t)s * nnb + f];
            const int orec = old_bc_of[(size_t)src[s] * nnb + f];
            if (rec < 0 || orec < 0) continue;
            BcRec &r = c->bc[rec];
            r.type = old_bc[orec].type;
            if (old_bc[orec].voff >= 0) {
                const int m = dim == 2 ? c->slot_nc[s] : c->slot_nc[s] * c->slot_nc[s];
                r.voff = (int)c->bc_vals.size();
                c->bc_vals.insert(c->bc_vals.end(), old_bc_vals.begin() + old_bc[orec].voff,
                                  old_bc_vals.begin() + old_bc[orec].voff + m);
            }
        }
    }
    c->bc_dirty = true;
    if (int e = upload_bc(c)) return e;
    // tile shapes and pass fusion of the levels that remain
    for (Level &L : c->levels) {
        auto it = old_level.find(L.lvl);
        if (it == old_level.end() || it->second.nc != L.nc) continue;
        for (int k = 0; k < 2; k++) {
            L.tile[k] = it->second.tile[k];
            L.pairs[k] = it->second.pairs[k];
        }
    }
    return 0;
}

int afg_update_tree(afg_ctx *c, int n_boxes, const int32_t *recs, double dr_base, const int32_t *tree_of,
                    int prolong_order, int n_vars, const int32_t *vars)
{
    return update_tree(c, n_boxes, recs, dr_base, tree_of, nullptr, prolong_order, n_vars, vars);
}

int afg_update_forest(afg_ctx *c, int n_boxes, const int32_t *recs, double dr_base, const int32_t *tree_of,
                      const int32_t *old_tree)
{
    if (!c) return AFG_ERR_ARG;
    if (!old_tree) return c->fail(AFG_ERR_ARG, "null predecessor list");
    return update_tree(c, n_boxes, recs, dr_base, tree_of, old_tree, -1, 0, nullptr);
}

int afg_set_bc(afg_ctx *c, int id, int nb, int bc_type, const double *b)
{
    // host tables only (no HIP call): a caller sets every physical face
    // before each solve; the table is sent when it changed
    if (!c) return AFG_ERR_ARG;
    if (!c->have_tree) return c->fail(AFG_ERR_STATE, "no tree set (call afg_set_tree first)");
    if (id < 1 || id > c->max_id || c->id2slot[id] < 0) return c->fail(AFG_ERR_ARG, "no box with id %d", id);
    if (nb < 1 || nb > c->nnb) return c->fail(AFG_ERR_ARG, "neighbour direction %d out of range", nb);
    if (bc_type != AFG_BC_DIRICHLET && bc_type != AFG_BC_NEUMANN)
        return c->fail(AFG_ERR_ARG, "fill_bc: unknown boundary condition %d", bc_type);
    const int s = c->id2slot[id];
    const int rec = c->bc_of[(size_t)s * c->nnb + nb - 1];
    if (rec < 0) return c->fail(AFG_ERR_ARG, "box %d direction %d is not a physical boundary", id, nb);
    BcRec &r = c->bc[rec];
    bool changed = r.type != bc_type;
    r.type = bc_type;
    const int n = c->slot_nc[s];
    const int m = c->dim == 2 ? n : n * n;
    bool zero = true;
    if (b)
        for (int k = 0; k < m; k++) zero = zero && b[k] == 0.0 && !std::signbit(b[k]);
    if (zero) {
        changed = changed || r.voff >= 0;
        r.voff = -1;
    } else {
        if (r.voff < 0) {
            r.voff = (int)c->bc_vals.size();
            c->bc_vals.resize(c->bc_vals.size() + m);
            changed = true;
        }
        changed = changed || std::memcmp(&c->bc_vals[r.voff], b, m * sizeof(double)) != 0;
        std::memcpy(&c->bc_vals[r.voff], b, m * sizeof(double));
    }
    if (changed) c->bc_dirty = true;
    return 0;
}

int afg_packed_size(const afg_ctx *cc, int n_ids, const int32_t *ids, int64_t *n_values)
{
    afg_ctx *c = const_cast<afg_ctx *>(cc);
    if (int e = check_ctx(c)) return e;
    int64_t n = 0;
    for (int k = 0; k < n_ids; k++) {
        const int id = ids[k];
        if (id < 1 || id > c->max_id || c->id2slot[id] < 0) return c->fail(AFG_ERR_ARG, "no box with id %d", id);
        n += box_len(c->dim, c->slot_nc[c->id2slot[id]]);
    }
    *n_values = n;
    return 0;
}

// Copy the contiguous runs of marked slots between device pool and staging.
static int copy_runs(afg_ctx *c, double *dev, const std::vector<char> &mark, bool to_device, int stride)
{
    int s = 0;
    while (s < c->n_slots) {
        if (!mark[s]) {
            s++;
            continue;
        }
        int e = s;
        while (e < c->n_slots && mark[e]) e++;
        const size_t off = (size_t)s * stride, len = (size_t)(e - s) * stride * sizeof(double);
        if (to_device)
            HIP_TRY(c, hipMemcpyAsync(dev + off, c->h_stage + off, len, hipMemcpyHostToDevice, c->stream));
        else
            HIP_TRY(c, hipMemcpyAsync(c->h_stage + off, dev + off, len, hipMemcpyDeviceToHost, c->stream));
        s = e;
    }
    return 0;
}

}  // extern "C"

// ---- transfers of cell data between host box arrays and the device pools ----
//
// A variable moves through the context's pinned staging buffer, which holds
// it in the pool's slot layout.  Host copies (box array <-> staging) run on
// several threads; the slots are cut into XFER_PIECES pieces of consecutive
// slots, and the device copy of a piece is queued as soon as its boxes are
// staged (upload) or the boxes of a piece are copied out as soon as its device
// copy has arrived (download), so that host copies and PCIe transfers overlap.

constexpr int XFER_PIECES = 8;

// threads for the host copies: the CPUs this process may run on, at most 16
static int host_threads()
{
    static const int n = [] {
        cpu_set_t set;
        CPU_ZERO(&set);
        int k = sched_getaffinity(0, sizeof(set), &set) == 0 ? CPU_COUNT(&set) : 0;
        if (k < 1) k = (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(k, 16));
    }();
    return n;
}

// The boxes of ids by piece: slot_of[k] and, per piece, the indices k whose
// slot is in it; every id must be a box of the tree, once.
struct XferPlan {
    std::vector<int> slot_of;
    std::vector<int> piece[XFER_PIECES];
    std::vector<char> mark;
};

static int xfer_plan(afg_ctx *c, int n_ids, const int32_t *ids, const void *const *boxes, XferPlan &P)
{
    if (n_ids < 0 || (n_ids > 0 && (!ids || !boxes))) return c->fail(AFG_ERR_ARG, "transfer: invalid box list");
    P.slot_of.resize(n_ids);
    P.mark.assign(c->n_slots, 0);
    for (int k = 0; k < n_ids; k++) {
        const int id = ids[k];
        if (id < 1 || id > c->max_id || c->id2slot[id] < 0) return c->fail(AFG_ERR_ARG, "no box with id %d", id);
        if (!boxes[k]) return c->fail(AFG_ERR_ARG, "transfer: no array for box %d", id);
        const int s = c->id2slot[id];
        if (P.mark[s]) return c->fail(AFG_ERR_ARG, "transfer: box %d listed twice", id);
        P.mark[s] = 1;
        P.slot_of[k] = s;
        P.piece[(int)((long long)s * XFER_PIECES / c->n_slots)].push_back(k);
    }
    return 0;
}

// the slot range [lo, hi) of piece p
static void piece_slots(const afg_ctx *c, int p, int &lo, int &hi)
{
    // smallest s with s * P / n_slots >= p
    lo = (int)(((long long)p * c->n_slots + XFER_PIECES - 1) / XFER_PIECES);
    hi = (int)(((long long)(p + 1) * c->n_slots + XFER_PIECES - 1) / XFER_PIECES);
}

// device copies of the contiguous runs of marked slots in [lo, hi)
static hipError_t copy_slot_runs(afg_ctx *c, double *dev, const std::vector<char> &mark, bool to_device, int lo,
                                 int hi)
{
    int s = lo;
    while (s < hi) {
        if (!mark[s]) {
            s++;
            continue;
        }
        int e = s;
        while (e < hi && mark[e]) e++;
        const size_t off = (size_t)s * c->stride, len = (size_t)(e - s) * c->stride * sizeof(double);
        const hipError_t r = to_device
                                 ? hipMemcpyAsync(dev + off, c->h_stage + off, len, hipMemcpyHostToDevice, c->stream)
                                 : hipMemcpyAsync(c->h_stage + off, dev + off, len, hipMemcpyDeviceToHost, c->stream);
        if (r != hipSuccess) return r;
        s = e;
    }
    return hipSuccess;
}

// f(t, p, i) on T threads: thread t takes its share of the indices i of every
// piece p in turn; after(p) runs once per piece, on the thread that finished
// the piece last (upload), or before(p) on every thread before its share of p
template <typename Before, typename F, typename After>
static void xfer_threads(const XferPlan &P, Before before, F f, After after)
{
    int total = 0;
    for (const auto &pc : P.piece) total += (int)pc.size();
    const int T = std::max(1, std::min(host_threads(), total / 64));
    std::atomic<int> left[XFER_PIECES];
    for (int p = 0; p < XFER_PIECES; p++) left[p] = T;
    auto work = [&](int t) {
        for (int p = 0; p < XFER_PIECES; p++) {
            const int n = (int)P.piece[p].size();
            before(p);
            for (int i = (int)((long long)n * t / T); i < (int)((long long)n * (t + 1) / T); i++) f(P.piece[p][i]);
            if (left[p].fetch_sub(1) == 1) after(p);
        }
    };
    if (T == 1) {
        work(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
}

// the packed layout: box k's array at the sum of the lengths of boxes 0..k-1
static int packed_boxes(afg_ctx *c, int n_ids, const int32_t *ids, const double *packed, std::vector<const double *> &out)
{
    if (n_ids < 0 || (n_ids > 0 && (!ids || !packed))) return c->fail(AFG_ERR_ARG, "transfer: invalid box list");
    out.resize(n_ids);
    size_t pos = 0;
    for (int k = 0; k < n_ids; k++) {
        const int id = ids[k];
        if (id < 1 || id > c->max_id || c->id2slot[id] < 0) return c->fail(AFG_ERR_ARG, "no box with id %d", id);
        out[k] = packed + pos;
        pos += box_len(c->dim, c->slot_nc[c->id2slot[id]]);
    }
    return 0;
}

extern "C" {

int afg_upload_boxes(afg_ctx *c, int iv, int n_ids, const int32_t *ids, const double *const *boxes)
{
    if (int e = check_ctx(c)) return e;
    XferPlan P;
    if (int e = xfer_plan(c, n_ids, ids, (const void *const *)boxes, P)) return e;
    data_written(c);
    double *dev = nullptr;
    if (int e = ensure_var(c, iv, &dev)) return e;
    if (int e = ensure_stage(c, (size_t)c->n_slots * c->stride)) return e;
    if (int e_ = stream_sync(c)) return e_;   // the staging buffer may still be read by a device copy
    std::atomic<int> err{hipSuccess};
    xfer_threads(
        P, [](int) {},
        [&](int k) {
            const int s = P.slot_of[k];
            const int len = box_len(c->dim, c->slot_nc[s]);
            double *st = c->h_stage + (size_t)s * c->stride;
            std::memcpy(st, boxes[k], len * sizeof(double));
            if (len < c->stride) std::memset(st + len, 0, (c->stride - len) * sizeof(double));
        },
        [&](int p) {
            int lo, hi;
            piece_slots(c, p, lo, hi);
            const hipError_t r = copy_slot_runs(c, dev, P.mark, true, lo, hi);
            if (r != hipSuccess) err = r;
        });
    if (err != hipSuccess) return c->fail(AFG_ERR_HIP, "upload: %s", hipGetErrorString((hipError_t)err.load()));
    return 0;
}

int afg_download_boxes(afg_ctx *c, int iv, int n_ids, const int32_t *ids, double *const *boxes)
{
    if (int e = check_ctx(c)) return e;
    XferPlan P;
    if (int e = xfer_plan(c, n_ids, ids, (const void *const *)boxes, P)) return e;
    double *dev = nullptr;
    if (int e = ensure_var(c, iv, &dev)) return e;
    if (int e = ensure_stage(c, (size_t)c->n_slots * c->stride)) return e;
    for (int p = 0; p < XFER_PIECES; p++) {
        if (!c->xfer_ev[p]) HIP_TRY(c, hipEventCreateWithFlags(&c->xfer_ev[p], hipEventDisableTiming));
        int lo, hi;
        piece_slots(c, p, lo, hi);
        HIP_TRY(c, copy_slot_runs(c, dev, P.mark, false, lo, hi));
        HIP_TRY(c, hipEventRecord(c->xfer_ev[p], c->stream));
    }
    std::atomic<int> err{hipSuccess};
    xfer_threads(
        P,
        [&](int p) {
            const hipError_t r = hipEventSynchronize(c->xfer_ev[p]);
            if (r != hipSuccess) err = r;
        },
        [&](int k) {
            if (err != hipSuccess) return;
            const int s = P.slot_of[k];
            std::memcpy(boxes[k], c->h_stage + (size_t)s * c->stride, box_len(c->dim, c->slot_nc[s]) * sizeof(double));
        },
        [](int) {});
    if (err != hipSuccess) return c->fail(AFG_ERR_HIP, "download: %s", hipGetErrorString((hipError_t)err.load()));
    return 0;
}

int afg_upload_cc(afg_ctx *c, int iv, int n_ids, const int32_t *ids, const double *packed)
{
    if (int e = check_ctx(c)) return e;
    std::vector<const double *> boxes;
    if (int e = packed_boxes(c, n_ids, ids, packed, boxes)) return e;
    return afg_upload_boxes(c, iv, n_ids, ids, boxes.data());
}

int afg_download_cc(afg_ctx *c, int iv, int n_ids, const int32_t *ids, double *packed)
{
    if (int e = check_ctx(c)) return e;
    std::vector<const double *> boxes;
    if (int e = packed_boxes(c, n_ids, ids, packed, boxes)) return e;
    return afg_download_boxes(c, iv, n_ids, ids, const_cast<double *const *>(boxes.data()));
}

static int face_len(int dim, int n) { return dim * (n + 1) * (dim == 3 ? n * n : n); }

int afg_upload_fc(afg_ctx *c, int ifc, int n_ids, const int32_t *ids, const double *packed)
{
    if (int e = check_ctx(c)) return e;
    double *dev = nullptr;
    if (int e = ensure_fc(c, ifc, &dev)) return e;
    if (int e = ensure_stage(c, (size_t)c->n_slots * c->fstride)) return e;
    if (int e_ = stream_sync(c)) return e_;   // staging may be in use
    std::vector<char> mark(c->n_slots, 0);
    size_t pos = 0;
    for (int k = 0; k < n_ids; k++) {
        const int id = ids[k];
        if (id < 1 || id > c->max_id || c->id2slot[id] < 0) return c->fail(AFG_ERR_ARG, "no box with id %d", id);
        const int s = c->id2slot[id];
        const int len = face_len(c->dim, c->slot_nc[s]);
        std::memcpy(c->h_stage + (size_t)s * c->fstride, packed + pos, len * sizeof(double));
        if (len < c->fstride) std::memset(c->h_stage + (size_t)s * c->fstride + len, 0, (c->fstride - len) * sizeof(double));
        pos += len;
        mark[s] = 1;
    }
    if (int e = copy_runs(c, dev, mark, true, c->fstride)) return e;
    if (int e_ = stream_sync(c)) return e_;
    return 0;
}

int afg_download_fc(afg_ctx *c, int ifc, int n_ids, const int32_t *ids, double *packed)
{
    if (int e = check_ctx(c)) return e;
    double *dev = nullptr;
    if (int e = ensure_fc(c, ifc, &dev)) return e;
    if (int e = ensure_stage(c, (size_t)c->n_slots * c->fstride)) return e;
    std::vector<char> mark(c->n_slots, 0);
    for (int k = 0; k < n_ids; k++) {
        const int id = ids[k];
        if (id < 1 || id > c->max_id || c->id2slot[id] < 0) return c->fail(AFG_ERR_ARG, "no box with id %d", id);
        mark[c->id2slot[id]] = 1;
    }
    if (int e = copy_runs(c, dev, mark, false, c->fstride)) return e;
    if (int e_ = stream_sync(c)) return e_;
    size_t pos = 0;
    for (int k = 0; k < n_ids; k++) {
        const int s = c->id2slot[ids[k]];
        const int len = face_len(c->dim, c->slot_nc[s]);
        std::memcpy(packed + pos, c->h_stage + (size_t)s * c->fstride, len * sizeof(double));
        pos += len;
    }
    return 0;
}

// The work records of a$D_consistent_fluxes: for every parent of a level
// >= 1 (the boxes below level 1 always have children) and face whose
// neighbour exists and has no children, the neighbour, the face and the
// children on that face.
static int build_flux_records(afg_ctx *c)
{
    if (c->flux_ready) return 0;
    const int dim = c->dim, nnb = c->nnb, nch = c->nch;
    std::vector<FluxRec> recs;
    const int nl = (int)c->levels.size();
    c->flux_off.assign(nl + 1, 0);
    std::vector<int32_t> children;
    for (int li = 0; li < nl; li++) {
        const Level &L = c->levels[li];
        c->flux_off[li] = (int)recs.size();
        if (L.lvl < 1 || li + 1 >= nl) continue;
        // children slots of this level's boxes: the next level's parent slots
        const Level &C = c->levels[li + 1];
        children.assign((size_t)L.n * nch, -1);
        for (int k = 0; k < C.n; k++) {
            const int ps = c->slot_par[C.slot0 + k] - L.slot0;
            const int cs = C.slot0 + k;
            int ci = 0;
            for (int d = 0; d < dim; d++) ci |= ((c->slot_ix[(size_t)cs * 3 + d] - 1) & 1) << d;   // a$D_ix_to_ichild
            children[(size_t)ps * nch + ci] = cs;
        }
        for (int b = 0; b < L.n; b++) {
            if (children[(size_t)b * nch] < 0) continue;   // a leaf
            const int s = L.slot0 + b;
            for (int f = 0; f < nnb; f++) {
                const int nid = c->h_nb[(size_t)s * nnb + f];
                if (nid <= 0) continue;
                const int ns = c->id2slot[nid];
                if (children[(size_t)(ns - L.slot0) * nch] >= 0) continue;   // has children
                FluxRec r;
                r.nb = ns;
                r.f = f;
                const int d = f >> 1, high = f & 1;
                const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
                for (int q = 0; q < 4; q++) r.child[q] = -1;
                for (int q = 0; q < (1 << (dim - 1)); q++) {
                    const int ci = (high << d) | ((q & 1) << ta) | (dim == 3 ? ((q >> 1) & 1) << tb : 0);
                    r.child[q] = children[(size_t)b * nch + ci];
                }
                recs.push_back(r);
            }
        }
    }
    c->flux_off[nl] = (int)recs.size();
    if (!recs.empty()) {
        HIP_TRY(c, ctx_malloc(c, &c->d_flux, recs.size() * sizeof(FluxRec)));
        HIP_H2D(c, c->d_flux, recs.data(), recs.size() * sizeof(FluxRec));
    }
    c->flux_ready = true;
    return 0;
}

int afg_consistent_fluxes(afg_ctx *c, int n_f, const int32_t *f_ixs)
{
    if (int e = check_ctx(c)) return e;
    if (n_f < 0 || (n_f > 0 && !f_ixs)) return c->fail(AFG_ERR_ARG, "invalid face-variable list");
    std::vector<double *> pools(n_f);
    for (int k = 0; k < n_f; k++)
        if (int e = ensure_fc(c, f_ixs[k], &pools[k])) return e;
    if (int e = build_flux_records(c)) return e;
    for (size_t li = 0; li + 1 < c->flux_off.size(); li++) {
        const int n = c->flux_off[li + 1] - c->flux_off[li];
        for (double *p : pools)
            launch_flux(c->dim, c->stream, c->d_flux + c->flux_off[li], n, c->levels[li].nc, c->fstride, p,
                        c->cyl ? c->d_cyl_rmin : nullptr, c->levels[li].dr);
    }
    return launched(c, "afg_consistent_fluxes");
}

int afg_copy_cc(afg_ctx *c, int iv_from, int iv_to)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    double *a = nullptr, *b = nullptr;
    if (int e = ensure_var(c, iv_from, &a)) return e;
    if (int e = ensure_var(c, iv_to, &b)) return e;
    const DevTree T = dev_tree(c);
    for (const Level &L : c->levels) launch_copy_boxes(c->dim, c->stream, T, L.slot0, L.n, L.nc, a, b);
    return launched(c, "afg_copy_cc");
}

int afg_set_cc(afg_ctx *c, int iv, double value)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    double *a = nullptr;
    if (int e = ensure_var(c, iv, &a)) return e;
    const DevTree T = dev_tree(c);
    for (const Level &L : c->levels) launch_fill_value(c->dim, c->stream, T, L.slot0, L.n, L.nc, a, value);
    return launched(c, "afg_set_cc");
}

int afg_mg_init(afg_mg_opts *o)
{
    if (!o) return AFG_ERR_ARG;
    // mg$D_init_mg (m_aX_mg.f90:130-154)
    if (o->i_phi < 1 || o->i_tmp < 1 || o->i_rhs < 1) {
        g_create_error = o->i_phi < 1 ? "mg_init_mg: i_phi not set" : o->i_tmp < 1 ? "mg_init_mg: i_tmp not set" : "mg_init_mg: i_rhs not set";
        return AFG_ERR_ARG;
    }
    if (o->op == AFG_OP_AUTO && o->i_eps != -1 && o->i_eps < 1) {
        g_create_error = "mg_init_mg: invalid i_eps";
        return AFG_ERR_ARG;
    }
    // level set and boundary value: 0 or -1 is unset (zero-filled reserved words of older callers)
    if (o->i_lsf == 0) o->i_lsf = -1;
    if (o->i_bval == 0) o->i_bval = -1;
    if (o->i_lsf * o->i_bval < 0 || o->i_lsf < -1 || o->i_bval < -1) {
        g_create_error = "mg_init_mg: you have to set both i_lsf and i_bval";
        return AFG_ERR_ARG;
    }
    if (o->n_cycle_down < 0) o->n_cycle_down = 2;
    if (o->n_cycle_up < 0) o->n_cycle_up = 2;
    if (o->n_cycle_base < 0) o->n_cycle_base = 2;
    o->initialized = 1;
    return 0;
}

int afg_levels(const afg_ctx *cc, int *min_lvl, int *highest_lvl)
{
    afg_ctx *c = const_cast<afg_ctx *>(cc);
    if (int e = check_ctx(c)) return e;
    if (min_lvl) *min_lvl = c->min_lvl;
    if (highest_lvl) *highest_lvl = c->highest_lvl;
    return 0;
}

int afg_level_ids(const afg_ctx *cc, int lvl, int n, int32_t *ids, int *n_out)
{
    afg_ctx *c = const_cast<afg_ctx *>(cc);
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, false)) return e;
    const Level &L = c->levels[lvl - c->min_lvl];
    if (n_out) *n_out = L.n;
    if (!ids) return 0;
    if (n < L.n) return c->fail(AFG_ERR_ARG, "level %d has %d boxes, room for %d", lvl, L.n, n);
    for (int b = 0; b < L.n; b++) ids[b] = c->slot2id[L.slot0 + b];
    return 0;
}

int afg_fas_vcycle(afg_ctx *c, const afg_mg_opts *o, int highest_lvl, int set_residual)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, highest_lvl, false)) return e;
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    if (int e = set_tags(m)) return e;
    if (int e = run_vcycle(m, highest_lvl, set_residual ? 1 : 0, c->min_lvl, c->highest_lvl + 1)) return e;
    return launched(c, "afg_fas_vcycle");
}

int afg_fas_fmg(afg_ctx *c, const afg_mg_opts *o, int set_residual, int have_guess)
{
    if (int e = check_ctx(c)) return e;
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    if (int e = set_tags(m)) return e;
    if (int e = fas_fmg(m, set_residual, have_guess)) return e;
    return launched(c, "afg_fas_fmg");
}

// ---- FP32 stages and the mixed-precision solve ----

int afg_fp32_from_cc(afg_ctx *c, int iv, int iv32)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    double *a = nullptr;
    float *b = nullptr;
    if (int e = ensure_var(c, iv, &a)) return e;
    if (int e = ensure_var(c, iv32, &b)) return e;
    launch_to_fp32(c->stream, (long long)c->n_slots * c->stride, a, b);
    return launched(c, "afg_fp32_from_cc");
}

int afg_fp32_to_cc(afg_ctx *c, int iv32, int iv, int add)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    double *a = nullptr;
    float *b = nullptr;
    if (int e = ensure_var(c, iv, &a)) return e;
    if (int e = ensure_var(c, iv32, &b)) return e;
    launch_from_fp32(c->stream, (long long)c->n_slots * c->stride, b, a, add ? 1 : 0);
    return launched(c, "afg_fp32_to_cc");
}

// FP32 solves need the tags of the auto operators, set from the FP64 eps
static int fp32_mg(afg_ctx *c, const afg_mg_opts *o, MgT<float> *m32)
{
    Mg m64;
    if (int e = make_mg(c, o, &m64)) return e;
    set_tags(m64);
    return make_mg(c, o, m32);
}

int afg_fas_fmg_fp32(afg_ctx *c, const afg_mg_opts *o, int set_residual, int have_guess)
{
    if (int e = check_ctx(c)) return e;
    MgT<float> m;
    if (int e = fp32_mg(c, o, &m)) return e;
    if (int e = fas_fmg(m, set_residual, have_guess)) return e;
    return launched(c, "afg_fas_fmg_fp32");
}

int afg_fas_vcycle_fp32(afg_ctx *c, const afg_mg_opts *o, int highest_lvl, int set_residual)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, highest_lvl, false)) return e;
    MgT<float> m;
    if (int e = fp32_mg(c, o, &m)) return e;
    if (int e = run_vcycle(m, highest_lvl, set_residual ? 1 : 0, c->min_lvl, c->highest_lvl + 1)) return e;
    return launched(c, "afg_fas_vcycle_fp32");
}

static int tree_extreme(afg_ctx *c, int iv, int include_parents, int want_max, double *out);

// max |variable iv| over leaf interiors
static int leaf_max_abs(afg_ctx *c, int iv, double *out)
{
    double hi = 0, lo = 0;
    if (int e = tree_extreme(c, iv, 0, 1, &hi)) return e;
    if (int e = tree_extreme(c, iv, 0, 0, &lo)) return e;
    *out = std::max(std::fabs(hi), std::fabs(lo));
    return 0;
}

// Per tree: max |variable iv| over its leaf interiors, into out[0..n_trees).
static int forest_norms(afg_ctx *c, int iv, double *out)
{
    double *v = nullptr;
    if (int e = ensure_var(c, iv, &v)) return e;
    launch_tree_maxabs(c->dim, c->stream, c->stride, c->nc, c->d_tree_leaves, c->d_chunk, c->n_chunks,
                       c->d_tree_chunk, c->n_trees, v, c->d_chunk_max, c->d_tree_out);
    if (int e = launched(c, "forest norms")) return e;
    HIP_TRY(c, hipMemcpyAsync(c->h_tree_out, c->d_tree_out, c->n_trees * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    if (int e_ = stream_sync(c)) return e_;
    std::memcpy(out, c->h_tree_out, c->n_trees * sizeof(double));
    return 0;
}

// The FP64 residual of the mixed-precision solve: phi of every parent set to
// the restriction of its children (where the FAS iteration converges to, its
// corrections vanishing), then the ghost fill from the lowest level up
// (refinement-boundary ghosts read the parent's) and tmp = rhs - A phi.
// The FP64 residual on every level (parents' phi restricted from their
// children first); with f, its FP32 image goes straight into f's rhs and
// colour-split rhs, which the next correction then takes without converting
// tmp (mixed_r32: cleared by any write of cell data in between).
static void mixed_residual(Mg &m, MgT<float> *f = nullptr)
{
    afg_ctx *c = m.c;
    for (int lvl = c->highest_lvl; lvl > c->min_lvl; lvl--)
        launch_restrict(c->dim, c->stream, m.T, level_args(c, lvl), m.V.phi, m.V.phi);
    const bool to32 = f && c->res_upper > c->highest_lvl;
    for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++) {
        fill_gc(m, m.V.phi, lvl);
        if (to32)
            launch_residual32(c->dim, c->stream, m.T, level_args(c, lvl), level_vars(m, lvl), f->V.rhs, f->V.rhs_rb,
                              f->V.rb_stride);
        else
            residual_level(m, lvl);
    }
    if (to32) {
        c->mixed_r32 = true;
        c->mixed_r32_of[0] = m.V.tmp;
        c->mixed_r32_of[1] = f->V.rhs;
        if (f->V.rhs_rb) {
            c->rb_valid[1] = true;
            c->rb_var[1] = f->o->i_rhs;
        }
    }
}

int afg_mixed_begin(afg_ctx *c, const afg_mg_opts *o, int have_guess)
{
    if (int e = check_ctx(c)) return e;
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    if (int e = set_tags(m)) return e;
    MgT<float> f;
    if (int e = make_mg(c, o, &f)) return e;
    if (m.V.op_auto)
        launch_to_fp32(c->stream, (long long)c->n_slots * c->stride, m.V.eps, const_cast<float *>(f.V.eps));
    if (!have_guess)
        for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++) {
            const LevelArgs L = level_args(c, lvl);
            launch_fill_value(c->dim, c->stream, m.T, L.slot0, L.n, L.nc, m.V.phi, 0.0);
        }
    mixed_residual(m, &f);
    c->mixed_fmg = c->mixed_first_fmg;
    c->mixed_ready = true;
    return launched(c, "afg_mixed_begin");
}

int afg_mixed_correct(afg_ctx *c, const afg_mg_opts *o, int cycles)
{
    if (int e = check_ctx(c)) return e;
    if (!c->mixed_ready) return c->fail(AFG_ERR_STATE, "afg_mixed_begin has not been called");
    if (cycles < 1) return c->fail(AFG_ERR_ARG, "cycles per correction must be >= 1");
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    MgT<float> f;
    if (int e = make_mg(c, o, &f)) return e;
    const long long n = (long long)c->n_slots * c->stride;
    // correction e from A e = r in FP32: `cycles` V-cycles from e = 0 (the
    // first of the first correction an FMG), then phi += e
    if (!c->mixed_r32 || c->mixed_r32_of[0] != m.V.tmp || c->mixed_r32_of[1] != f.V.rhs) {
        launch_to_fp32(c->stream, n, m.V.tmp, f.V.rhs);
        if (f.V.rhs_rb) refresh_rhs_rb(c, f.T, (const float *)f.V.rhs, f.V.rhs_rb, o->i_rhs);
    }
    c->mixed_r32 = false;   // the V-cycles write rhs (tau on the parents)
    HIP_TRY(c, hipMemsetAsync(f.V.phi, 0, n * sizeof(float), c->stream));
    HIP_TRY(c, hipMemsetAsync(f.V.tmp, 0, n * sizeof(float), c->stream));
    for (int k = 0; k < cycles; k++) {
        if (c->mixed_fmg) {
            if (int e = fas_fmg(f, 0, 0)) return e;
            c->mixed_fmg = false;
        } else {
            if (int e = run_vcycle(f, c->highest_lvl, 0, c->min_lvl, c->highest_lvl + 1)) return e;
        }
    }
    launch_from_fp32(c->stream, n, f.V.phi, m.V.phi, 1);
    mixed_residual(m, &f);
    return launched(c, "afg_mixed_correct");
}

int afg_mixed_fp64_base(afg_ctx *c, int enabled)
{
    if (!c) return AFG_ERR_ARG;
    if (c->stream) if (int e_ = stream_sync(c)) return e_;
    c->mixed_base64 = enabled != 0;
    for (auto &g : c->graphs) destroy_graph(g.second);   // the flag is a kernel argument
    c->graphs.clear();
    return 0;
}

int afg_mixed_fp64_tau(afg_ctx *c, int enabled)
{
    if (!c) return AFG_ERR_ARG;
    if (c->stream) if (int e_ = stream_sync(c)) return e_;
    c->mixed_tau64 = enabled != 0;
    for (auto &g : c->graphs) destroy_graph(g.second);   // the flag selects kernels
    c->graphs.clear();
    return 0;
}

int afg_mixed_first_fmg(afg_ctx *c, int enabled)
{
    if (!c) return AFG_ERR_ARG;
    c->mixed_first_fmg = enabled != 0;
    return 0;
}

int afg_mixed_solve(afg_ctx *c, const afg_mg_opts *o, int have_guess, double tol, int max_iter,
                    int cycles, int *iterations, double *rel_residual)
{
    if (int e = check_ctx(c)) return e;
    if (!(tol > 0) || max_iter < 0) return c->fail(AFG_ERR_ARG, "tol must be > 0 and max_iter >= 0");
    if (cycles < 1) return c->fail(AFG_ERR_ARG, "cycles per correction must be >= 1");
    if (o && (o->i_rhs < 1 || o->i_rhs > c->n_var)) return c->fail(AFG_ERR_ARG, "bad i_rhs");
    if (int e = afg_mixed_begin(c, o, have_guess)) return e;
    const int nt = c->n_trees;
    std::vector<double> rhs_norm(nt), res(nt);
    if (int e = forest_norms(c, o->i_rhs, rhs_norm.data())) return e;
    int it = 0;
    double worst = 0;
    for (;; it++) {
        if (int e = forest_norms(c, o->i_tmp, res.data())) return e;
        bool done = true;
        worst = 0;
        for (int t = 0; t < nt; t++) {
            done = done && res[t] <= tol * rhs_norm[t];
            worst = std::max(worst, rhs_norm[t] > 0 ? res[t] / rhs_norm[t] : res[t]);
        }
        if (done || it >= max_iter) break;
        if (int e = afg_mixed_correct(c, o, cycles)) return e;
    }
    if (iterations) *iterations = it;
    if (rel_residual) *rel_residual = worst;
    return 0;
}

int afg_tree_extreme_noalloc(afg_ctx *c, int iv, int include_parents, int want_max, double *d_out)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_var(c, iv)) return e;
    if (!d_out) return c->fail(AFG_ERR_ARG, "null output");
    if (c->capturing) return c->fail(AFG_ERR_STATE, "afg_tree_extreme_noalloc inside the library's own capture");
    const double *v = c->d_var[iv - 1];
    if (!v) return c->fail(AFG_ERR_STATE, "variable %d has no device data (allocating it is not allowed here)", iv);
    if (!c->scratch) return c->fail(AFG_ERR_STATE, "no scratch array on device %d", c->device);
    int used = 0;
    for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++)
        used += extreme_blocks(c->dim, list_args(c, lvl, include_parents ? 2 : 1));
    if (used > AFG_NOALLOC_MAX_PARTIALS)
        return c->fail(AFG_ERR_UNSUPPORTED, "the tree needs %d partial results; the allocation-free reduction holds %d",
                       used, AFG_NOALLOC_MAX_PARTIALS);
    const DevTree T = dev_tree(c);
    int off = 0;
    for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++) {
        const ListArgs P = list_args(c, lvl, include_parents ? 2 : 1);
        launch_extreme_list(c->dim, c->stream, T, P, v, want_max ? 1 : 0, c->scratch + off);
        off += extreme_blocks(c->dim, P);
    }
    // every partial is written before it is read: nothing to clear; with no
    // partials the reduction writes the identity, as afg_tree_max_cc returns
    launch_reduce_partials(c->stream, c->scratch, used, want_max ? 1 : 0, d_out);
    return launched(c, "afg_tree_extreme_noalloc");
}

int afg_forest_max_abs(afg_ctx *c, int iv, int n_trees, double *out)
{
    if (int e = check_ctx(c)) return e;
    if (!out) return c->fail(AFG_ERR_ARG, "null output");
    if (n_trees != c->n_trees) return c->fail(AFG_ERR_ARG, "the forest has %d trees, not %d", c->n_trees, n_trees);
    return forest_norms(c, iv, out);
}

int afg_forest_trees(const afg_ctx *c, int *n_trees, int32_t *highest)
{
    if (!c || !c->have_tree || !n_trees) return AFG_ERR_ARG;
    *n_trees = c->n_trees;
    if (highest)
        for (int t = 0; t < c->n_trees; t++) highest[t] = c->tree_hi[t];
    return 0;
}

static int tree_extreme(afg_ctx *c, int iv, int include_parents, int want_max, double *out)
{
    if (int e = check_ctx(c)) return e;
    if (!out) return c->fail(AFG_ERR_ARG, "null output");
    double *v = nullptr;
    if (int e = ensure_var(c, iv, &v)) return e;
    const DevTree T = dev_tree(c);
    int used = 0;
    for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++) {
        const ListArgs P = list_args(c, lvl, include_parents ? 2 : 1);
        launch_extreme_list(c->dim, c->stream, T, P, v, want_max, c->d_partial + used);
        used += extreme_blocks(c->dim, P);
    }
    if (used == 0) {
        *out = want_max ? -1.7976931348623157e308 : 1.7976931348623157e308;
        return 0;
    }
    launch_reduce_partials(c->stream, c->d_partial, used, want_max, c->d_out);
    if (int e = launched(c, "tree extreme")) return e;
    HIP_TRY(c, hipMemcpyAsync(c->h_out, c->d_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (int e_ = stream_sync(c)) return e_;
    *out = c->h_out[0];
    return 0;
}

int afg_tree_max_cc(afg_ctx *c, int iv, int include_parents, double *out)
{
    return tree_extreme(c, iv, include_parents, 1, out);
}

int afg_tree_min_cc(afg_ctx *c, int iv, int include_parents, double *out)
{
    return tree_extreme(c, iv, include_parents, 0, out);
}

int afg_fill_gc(afg_ctx *c, int iv, int lvl)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, false)) return e;
    double *v = nullptr;
    if (int e = ensure_var(c, iv, &v)) return e;
    if (int e = upload_bc(c)) return e;
    launch_fill_gc(c->dim, c->stream, dev_tree(c), level_args(c, lvl), v);
    return launched(c, "afg_fill_gc");
}

int afg_gsrb_level(afg_ctx *c, const afg_mg_opts *o, int lvl, int n_cycle)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, false)) return e;
    if (n_cycle < 0) return c->fail(AFG_ERR_ARG, "n_cycle < 0");
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    if (int e = set_tags(m)) return e;
    gsrb_level(m, lvl, n_cycle);
    return launched(c, "afg_gsrb_level");
}

int afg_residual_level(afg_ctx *c, const afg_mg_opts *o, int lvl)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, false)) return e;
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    if (int e = set_tags(m)) return e;
    residual_level(m, lvl);
    return launched(c, "afg_residual_level");
}

int afg_restrict_level(afg_ctx *c, int iv, int lvl)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    if (int e = check_level(c, lvl, true)) return e;
    double *v = nullptr;
    if (int e = ensure_var(c, iv, &v)) return e;
    launch_restrict(c->dim, c->stream, dev_tree(c), level_args(c, lvl), v, v);
    return launched(c, "afg_restrict_level");
}

int afg_prolong2_level(afg_ctx *c, int iv, int iv_to, int lvl, int add)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    if (int e = check_level(c, lvl, true)) return e;
    double *a = nullptr, *b = nullptr;
    if (int e = ensure_var(c, iv, &a)) return e;
    if (int e = ensure_var(c, iv_to, &b)) return e;
    if (a == b) return c->fail(AFG_ERR_ARG, "prolong2: source and destination must differ");
    launch_prolong2(c->dim, c->stream, dev_tree(c), level_args(c, lvl), a, b, add ? 1 : 0);
    return launched(c, "afg_prolong2_level");
}

int afg_correct_children(afg_ctx *c, const afg_mg_opts *o, int lvl)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, true)) return e;
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    if (int e = set_tags(m)) return e;
    // correct_children (m_aX_mg.f90:368-395): tmp = phi - tmp on the parents,
    // then the prolongation of tmp into the children
    launch_diff(c->dim, c->stream, m.T, list_args(c, lvl - 1, 0), m.V);
    launch_correct(c->dim, c->stream, m.T, level_args(c, lvl), m.V, 0);
    return launched(c, "afg_correct_children");
}

int afg_update_coarse(afg_ctx *c, const afg_mg_opts *o, int lvl)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, true)) return e;
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    if (int e = set_tags(m)) return e;
    update_coarse(m, lvl, 0);
    return launched(c, "afg_update_coarse");
}

// The per-slot tables of a cylindrical tree: (ix(1) - 1) * nc_l and
// box%r_min(1), computed as the reference does: r_base + (ix - 1) dr_base
// n_cell on levels <= 1 (m_aX_core.f90:195-196), the parent's plus
// 0.5 dr_p dix n_cell_p for a child (:774-775); slots are in level order, so
// parents come first.
static int build_cyl(afg_ctx *c)
{
    ctx_free(c, c->d_cyl_ioff);
    ctx_free(c, c->d_cyl_rmin);
    c->d_cyl_ioff = nullptr;
    c->d_cyl_rmin = nullptr;
    if (!c->cyl || !c->have_tree) return 0;
    const int ns = c->n_slots;
    std::vector<int32_t> ioff(ns);
    std::vector<double> rmin(ns);
    for (int s = 0; s < ns; s++) {
        const int ix = c->slot_ix[(size_t)s * 3], lvl = c->slot_lvl[s];
        ioff[s] = (ix - 1) * c->slot_nc[s];
        if (lvl <= 1) {
            rmin[s] = c->r_base + (ix - 1) * c->dr_base * c->nc;
        } else {
            const int p = c->slot_par[s];
            const double dr_p = c->levels[lvl - 1 - c->min_lvl].dr;
            rmin[s] = rmin[p] + 0.5 * dr_p * ((ix - 1) & 1) * c->slot_nc[p];
        }
    }
    HIP_TRY(c, ctx_malloc(c, &c->d_cyl_ioff, ns * sizeof(int32_t)));
    HIP_TRY(c, ctx_malloc(c, &c->d_cyl_rmin, ns * sizeof(double)));
    HIP_H2D(c, c->d_cyl_ioff, ioff.data(), ns * sizeof(int32_t));
    HIP_H2D(c, c->d_cyl_rmin, rmin.data(), ns * sizeof(double));
    return 0;
}

static int check_gc_args(afg_ctx *c, int rb, int bc)
{
    if (rb < RB_MG || rb > RB_INTERP_LIM) return c->fail(AFG_ERR_ARG, "unknown refinement-boundary method %d", rb);
    if (bc != 0 && bc != AFG_BC_DIRICHLET && bc != AFG_BC_NEUMANN)
        return c->fail(AFG_ERR_ARG, "unknown boundary condition %d", bc);
    if ((rb == RB_INTERP || rb == RB_INTERP_LIM) && !c->rb_parent_nb)
        return c->fail(AFG_ERR_ARG, "gc_interp: a refinement boundary without a neighbour of the parent");
    return 0;
}

int afg_gc_level(afg_ctx *c, int iv, int lvl, int rb, int bc)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, false)) return e;
    if (int e = check_gc_args(c, rb, bc)) return e;
    double *v = nullptr;
    if (int e = ensure_var(c, iv, &v)) return e;
    if (int e = upload_bc(c)) return e;
    launch_gc(c->dim, c->stream, dev_tree(c), level_args(c, lvl), v, rb, bc);
    return launched(c, "afg_gc_level");
}

int afg_gc_tree(afg_ctx *c, int iv, int rb, int bc)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_gc_args(c, rb, bc)) return e;
    double *v = nullptr;
    if (int e = ensure_var(c, iv, &v)) return e;
    if (int e = upload_bc(c)) return e;
    const DevTree T = dev_tree(c);
    for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++)
        launch_gc(c->dim, c->stream, T, level_args(c, lvl), v, rb, bc);
    return launched(c, "afg_gc_tree");
}

int afg_gc2_boxes(afg_ctx *c, int iv, int n_ids, const int32_t *ids, double *packed)
{
    if (int e = check_ctx(c)) return e;
    if (n_ids < 0 || (n_ids > 0 && (!ids || !packed))) return c->fail(AFG_ERR_ARG, "gc2: invalid box list");
    for (int k = 0; k < n_ids; k++)
        if (ids[k] < 1 || ids[k] > c->max_id || c->id2slot[ids[k]] < 0)
            return c->fail(AFG_ERR_ARG, "no box with id %d", ids[k]);
    if (!c->rb_parent_nb)
        return c->fail(AFG_ERR_ARG, "gc2_prolong1: a refinement boundary without a neighbour of the parent");
    double *v = nullptr;
    if (int e = ensure_var(c, iv, &v)) return e;
    const int gstride = 2 * c->dim * (c->dim == 2 ? c->nc : c->nc * c->nc);
    const size_t n = (size_t)c->n_slots * gstride;
    double *d_out = nullptr;
    HIP_TRY(c, ctx_malloc(c, &d_out, n * sizeof(double)));
    const DevTree T = dev_tree(c);
    for (int lvl = c->min_lvl; lvl <= c->highest_lvl; lvl++)
        launch_gc2(c->dim, c->stream, T, level_args(c, lvl), v, d_out, gstride);
    int e = launched(c, "afg_gc2_boxes");
    std::vector<double> h(n);
    if (!e) {
        hipError_t r = hipMemcpyAsync(h.data(), d_out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
        if (r != hipSuccess) e = c->fail(AFG_ERR_HIP, "afg_gc2_boxes: %s", hipGetErrorString(r));
    }
    ctx_free(c, d_out);
    if (e) return e;
    size_t pos = 0;
    for (int k = 0; k < n_ids; k++) {
        const int sl = c->id2slot[ids[k]], m = c->slot_nc[sl];
        const int len = 2 * c->dim * (c->dim == 2 ? m : m * m);
        std::memcpy(packed + pos, h.data() + (size_t)sl * gstride, len * sizeof(double));
        pos += len;
    }
    return 0;
}

int afg_restrict_tree(afg_ctx *c, int iv, int iv_to)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    double *a = nullptr, *b = nullptr;
    if (int e = ensure_var(c, iv, &a)) return e;
    if (int e = ensure_var(c, iv_to, &b)) return e;
    const DevTree T = dev_tree(c);
    for (int lvl = c->highest_lvl; lvl > c->min_lvl; lvl--)
        launch_restrict(c->dim, c->stream, T, level_args(c, lvl), (const double *)a, b);
    return launched(c, "afg_restrict_tree");
}

int afg_prolong_level(afg_ctx *c, int iv, int iv_to, int lvl, int order, int add)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    if (int e = check_level(c, lvl, true)) return e;
    if (order < 0 || order > 2) return c->fail(AFG_ERR_ARG, "prolongation order %d not in 0..2", order);
    double *a = nullptr, *b = nullptr;
    if (int e = ensure_var(c, iv, &a)) return e;
    if (int e = ensure_var(c, iv_to, &b)) return e;
    launch_prolong_level(c->dim, c->stream, dev_tree(c), level_args(c, lvl), a, b, order, add ? 1 : 0);
    return launched(c, "afg_prolong_level");
}

static int pool_arith(afg_ctx *c, int op, double a, int iv_x, double b, int iv_y)
{
    if (int e = check_ctx(c)) return e;
    data_written(c);
    double *x = nullptr, *y = nullptr;
    if (int e = ensure_var(c, iv_x, &x)) return e;
    if (int e = ensure_var(c, iv_y, &y)) return e;
    launch_arith(c->stream, (long long)c->n_slots * c->stride, op, a, b, x, y);
    return launched(c, "tree arithmetic");
}

int afg_add_cc(afg_ctx *c, int iv_from, int iv_to) { return pool_arith(c, ARITH_ADD, 0, iv_from, 0, iv_to); }
int afg_sub_cc(afg_ctx *c, int iv_from, int iv_to) { return pool_arith(c, ARITH_SUB, 0, iv_from, 0, iv_to); }
int afg_times_cc(afg_ctx *c, double a, int iv) { return pool_arith(c, ARITH_TIMES, a, iv, 0, iv); }
int afg_lincomb_cc(afg_ctx *c, double a, int iv_a, double b, int iv_b)
{
    return pool_arith(c, ARITH_LINCOMB, a, iv_a, b, iv_b);
}

int afg_set_bottom_cells(afg_ctx *c, int64_t cells, int64_t wave_cells)
{
    if (int e = check_ctx(c)) return e;
    if (cells < 0 || wave_cells < 0) return c->fail(AFG_ERR_ARG, "bottom cell limits must be >= 0");
    if (int e_ = stream_sync(c)) return e_;
    c->bottom_cells[0] = c->bottom_cells[1] = cells;
    c->bottom_wave_cells = wave_cells;
    return choose_bottom(c);
}

int afg_set_split_passes(afg_ctx *c, int enabled)
{
    if (int e = check_ctx(c)) return e;
    if (int e_ = stream_sync(c)) return e_;
    c->split_passes = enabled != 0;
    for (auto &g : c->graphs) destroy_graph(g.second);
    c->graphs.clear();
    return 0;
}

int afg_set_persist(afg_ctx *c, int wgs, int64_t cells)
{
    if (int e = check_ctx(c)) return e;
    if (wgs < 0 || wgs > PERSIST_MAX_RANKS)
        return c->fail(AFG_ERR_ARG, "persistent kernel workgroups %d outside 0..%d", wgs, PERSIST_MAX_RANKS);
    if (cells < 0) return c->fail(AFG_ERR_ARG, "persistent kernel cell limit must be >= 0");
    if (int e_ = stream_sync(c)) return e_;
    c->persist_wgs = wgs;
    c->persist_cells[0] = c->persist_cells[1] = cells;
    return choose_bottom(c);
}

int afg_persist_trace(afg_ctx *c, int cap, int64_t *ticks, int *n)
{
    if (int e = check_ctx(c)) return e;
    if (cap < 0) return c->fail(AFG_ERR_ARG, "trace capacity must be >= 0");
    if (int e_ = stream_sync(c)) return e_;
    if (ticks && n && c->d_ptrace) {
        std::vector<unsigned long long> h(c->ptrace_cap);
        HIP_TRY(c, hipMemcpyAsync(h.data(), c->d_ptrace, h.size() * sizeof(h[0]), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        int k = 0;
        while (k < (int)h.size() && h[k] != 0) {
            ticks[k] = (int64_t)h[k];
            k++;
        }
        *n = k;
    }
    if (cap != c->ptrace_cap) {
        ctx_free(c, c->d_ptrace);
        c->d_ptrace = nullptr;
        c->ptrace_cap = 0;
        if (cap > 0) {
            HIP_TRY(c, ctx_malloc(c, &c->d_ptrace, (size_t)cap * sizeof(unsigned long long)));
            c->ptrace_cap = cap;
        }
        for (auto &g : c->graphs) destroy_graph(g.second);
        c->graphs.clear();
    }
    if (c->d_ptrace) {
        HIP_TRY(c, hipMemsetAsync(c->d_ptrace, 0, (size_t)c->ptrace_cap * sizeof(unsigned long long), c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int afg_get_persist(const afg_ctx *c, int *wgs, int64_t *cells, int *top_lvl)
{
    if (!c) return AFG_ERR_ARG;
    if (wgs) *wgs = c->persist_wgs;
    if (cells) *cells = c->persist_cells[0];
    if (top_lvl) *top_lvl = c->persist_wgs > 0 ? c->persist_lvl[0] : -1000;
    return 0;
}

int afg_set_tile(afg_ctx *c, int lvl, int fp32, int shape)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, false)) return e;
    Level &L = c->levels[lvl - c->min_lvl];
    if (shape < 0 || shape >= tile_shapes(L.nc))
        return c->fail(AFG_ERR_ARG, "level %d (nc %d) has tile shapes 0..%d", lvl, L.nc, tile_shapes(L.nc) - 1);
    if (int e_ = stream_sync(c)) return e_;
    L.tile[fp32 != 0] = shape;
    for (auto &g : c->graphs) destroy_graph(g.second);
    c->graphs.clear();
    return 0;
}

int afg_get_tile(const afg_ctx *c, int lvl, int fp32, int *shape)
{
    if (!c || !c->have_tree || !shape || lvl < c->min_lvl || lvl > c->highest_lvl) return AFG_ERR_ARG;
    *shape = c->levels[lvl - c->min_lvl].tile[fp32 != 0];
    return 0;
}

int afg_autotune(afg_ctx *c, const afg_mg_opts *o, int fp32, int reps, double *ms)
{
    if (int e = check_ctx(c)) return e;
    if (reps < 1) return c->fail(AFG_ERR_ARG, "reps must be >= 1");
    if (c->capturing) return c->fail(AFG_ERR_STATE, "autotune inside a graph capture");
    if (int e_ = stream_sync(c)) return e_;
    if (ms)
        for (int i = 0; i < 6 * (int)c->levels.size() + 4; i++) ms[i] = -1;
    Mg m;
    if (int e = make_mg(c, o, &m)) return e;
    if (int e = set_tags(m)) return e;
    if (int e = fp32 ? tune_tiles<float>(c, o, reps, ms, m.V.phi) : tune_tiles<double>(c, o, reps, ms, m.V.phi))
        return e;
    double *bms = ms ? ms + 6 * (int)c->levels.size() : nullptr;
    if (int e = fp32 ? tune_bottom<float>(c, o, reps, bms) : tune_bottom<double>(c, o, reps, bms)) return e;
    return launched(c, "afg_autotune");
}

int afg_get_bottom_cells(const afg_ctx *c, int fp32, int64_t *cells)
{
    if (!c || !cells) return AFG_ERR_ARG;
    *cells = c->bottom_cells[fp32 != 0];
    return 0;
}

int afg_set_pass_pairs(afg_ctx *c, int enabled)
{
    if (int e = check_ctx(c)) return e;
    if (int e_ = stream_sync(c)) return e_;
    for (Level &L : c->levels) L.pairs[0] = L.pairs[1] = enabled != 0;
    for (auto &g : c->graphs) destroy_graph(g.second);
    c->graphs.clear();
    return 0;
}

int afg_set_level_pass_pairs(afg_ctx *c, int lvl, int fp32, int enabled)
{
    if (int e = check_ctx(c)) return e;
    if (int e = check_level(c, lvl, false)) return e;
    if (int e_ = stream_sync(c)) return e_;
    c->levels[lvl - c->min_lvl].pairs[fp32 != 0] = enabled != 0;
    for (auto &g : c->graphs) destroy_graph(g.second);
    c->graphs.clear();
    return 0;
}

int afg_get_pass_pairs(const afg_ctx *c, int lvl, int fp32, int *enabled)
{
    if (!c || !c->have_tree || !enabled || lvl < c->min_lvl || lvl > c->highest_lvl) return AFG_ERR_ARG;
    *enabled = c->levels[lvl - c->min_lvl].pairs[fp32 != 0] ? 1 : 0;
    return 0;
}

int afg_set_graphs(afg_ctx *c, int enabled)
{
    if (!c) return AFG_ERR_ARG;
    c->use_graphs = enabled != 0;
    return 0;
}

int afg_set_timing(afg_ctx *c, int enabled, int lvl)
{
    if (!c) return AFG_ERR_ARG;
    if (int e = collect_timing(c)) return e;
    // with the persistent kernel on, the timed level runs by level: levels may
    // move between the two paths, and only the per-level stages keep the
    // colour-split rhs current
    if (c->persist_wgs > 0 && (c->timing != (enabled != 0) || c->timing_lvl != lvl)) data_written(c);
    c->timing = enabled != 0;
    c->timing_lvl = lvl;
    if (c->timing) {
        c->t_launches = 0;
        c->t_ms = 0;
        c->t_bytes = 0;
    }
    return 0;
}

int afg_timing(afg_ctx *c, int64_t *n_launches, double *ms, double *bytes)
{
    if (!c) return AFG_ERR_ARG;
    if (int e = collect_timing(c)) return e;
    if (n_launches) *n_launches = c->t_launches;
    if (ms) *ms = c->t_ms;
    if (bytes) *bytes = c->t_bytes;
    return 0;
}

int afg_vcycle_cell_updates(const afg_ctx *cc, const afg_mg_opts *o, int highest_lvl, double *out)
{
    afg_ctx *c = const_cast<afg_ctx *>(cc);
    if (int e = check_ctx(c)) return e;
    if (!o || !out) return c->fail(AFG_ERR_ARG, "null argument");
    if (int e = check_level(c, highest_lvl, false)) return e;
    double n = 0;
    for (int lvl = c->min_lvl; lvl <= highest_lvl; lvl++) {
        const Level &L = c->levels[lvl - c->min_lvl];
        const double cells = (double)L.n * (c->dim == 2 ? L.nc * L.nc : L.nc * L.nc * L.nc);
        const int sweeps = lvl == c->min_lvl ? o->n_cycle_base : o->n_cycle_down + o->n_cycle_up;
        n += cells * sweeps;
    }
    *out = n;
    return 0;
}

int afg_device_bytes(const afg_ctx *cc, int64_t *out)
{
    afg_ctx *c = const_cast<afg_ctx *>(cc);
    if (!c || !out) return AFG_ERR_ARG;
    *out = c->dev_bytes;
    return 0;
}

}  // extern "C"
