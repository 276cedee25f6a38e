// afg_types.h -- plain data types shared by the host orchestration and the
// kernel bodies of the afivo_amd multigrid library (no HIP dependency).
#pragma once

#include <cstddef>
#include <cstdint>

namespace afg {

// Encoding of the per-slot neighbour table (the device form of
// box%neighbors, m_aX_t.f90:163):
//   >= 0        slot of the same-level neighbour
//   NB_REFINE   no neighbour at this level: refinement boundary (a5_no_box)
//   <= NB_BC0   physical boundary, boundary record index NB_BC0 - value
constexpr int32_t NB_REFINE = -1;
constexpr int32_t NB_BC0 = -2;

// Tags of mg$D_set_box_tag (m_aX_mg.f90:13-17, 667-706).
constexpr int8_t TAG_UNSET = 0;
constexpr int8_t TAG_NORMAL = 1;
constexpr int8_t TAG_LSF = 2;
constexpr int8_t TAG_CEPS = 3;
constexpr int8_t TAG_VEPS = 4;
constexpr int8_t TAG_CONFLICT = -1;   // both a level-set and an eps tag: an error

// Refinement-boundary methods of the general ghost fill (subr_rb of
// a$D_gc_tree, m_aX_gc.f90:28-69).
constexpr int RB_MG = 0;           // mg$D_sides_rb (m_aX_mg.f90:262-350)
constexpr int RB_PROLONG0 = 1;     // a$D_gc_prolong0 (m_aX_gc.f90:135-151)
constexpr int RB_INTERP = 2;       // a$D_gc_interp (m_aX_gc.f90:155-251)
constexpr int RB_INTERP_LIM = 3;   // a$D_gc_interp_lim (m_aX_gc.f90:254-358)

// Whole-pool arithmetic: y = y + x, y - x, a y, a x + b y
// (a$D_box_add_cc, _sub_cc, _times_cc, _lincomb_cc, m_aX_utils.f90:273-318).
constexpr int ARITH_ADD = 0, ARITH_SUB = 1, ARITH_TIMES = 2, ARITH_LINCOMB = 3;

// One work record of a$D_consistent_fluxes (m_aX_core.f90:811-943): the
// leaf neighbour `nb` of a parent across the parent's face f, and the
// parent's children on that face, indexed by their tangential offset bits
// (first tangential axis | second << 1).
struct FluxRec {
    int32_t nb, f;
    int32_t child[4];
};

// Device topology of one tree.  Slots of a level are contiguous and in Morton
// order; every box array of a variable occupies `stride` doubles at
// var + slot * stride and holds cc(0:nc_l+1, ...) in Fortran order.
struct DevTree {
    int stride;
    const int32_t *nb;        // [slot * 2*dim + face]
    const int32_t *parent;    // parent slot, -1 if none
    const int32_t *choff;     // child offset in the parent: x | y << 10 | z << 20
    const int32_t *bc_type;   // per boundary record
    const int32_t *bc_voff;   // offset of the boundary values, -1 for b = 0
    const double *bc_vals;
    int8_t *tag;              // per slot (auto operators)
    const int8_t *is_parent;  // per slot: the box has children
    // cylindrical coordinates (2D, x the radius; null when Cartesian): per
    // slot (ix(1) - 1) * nc_l and box%r_min(1)
    const int32_t *cyl_ioff;
    const double *cyl_rmin;
};

// One level: slots [slot0, slot0 + n), nc cells per dimension, cell size dr;
// pnc is the cell count of the parent level.
// Number of tile shapes of the tiled colour pass for a cell count: nc 8 and
// 16 have three, the others one.
constexpr int tile_shapes(int nc) { return nc == 8 || nc == 16 ? 3 : 1; }

struct LevelArgs {
    int slot0, n, nc, pnc;
    int tile;           // shape of the tiled colour pass (TileShape), per level and precision
    double dr;          // box%dr of the level
    double dx2;         // dr * dr  (dr**2, m_aX_mg.f90:730)
    double inv_dr_sq;   // 1 / (dr * dr)  (m_aX_mg.f90:772)
};

// A list of slots (parents or leaves of a level), all with nc cells.
struct ListArgs {
    const int32_t *slots;
    int n, nc;
    double dr, dx2, inv_dr_sq;
};

// Variables of the multigrid options, as device pointers (R: double, or
// float for the FP32 stages of the mixed-precision solve).
template <typename R>
struct MgVarsT {
    R *phi, *rhs, *tmp;
    const R *eps;   // nullptr unless the auto operators are used
    int op_auto;
    // level set and boundary value (nullptr unless used); lsf_mode: 0 none,
    // 1 boxes tagged TAG_LSF, 2 every box; rstr_mode: 0 rstr_lpl, 1 by the
    // child's tag, 2 rstr_lpllsf for every box
    const R *lsf = nullptr, *bval = nullptr;
    int lsf_mode = 0, rstr_mode = 0;
    // rhs by colour, without ghosts (nullptr: not kept; 3D only): slot s holds at
    // s * rb_stride the interior values of colour 1, then of colour 2, each in
    // the order of colour_cell; the tiled colour passes read it in place of
    // rhs (a pass then reads half of the interior instead of every cache
    // line of the box), and the stages that write rhs keep it current
    R *rhs_rb = nullptr;
    int rb_stride = 0;
    // phi by colour in the same layout, the working copy of the colour passes
    // of a relaxation block (3D tiles; nullptr: not used)
    R *phi_rb = nullptr;
    // FP32 stages: update_coarse's arithmetic (the fine residual, the
    // restriction sums and A phi + tmp on the parents, which form the FAS
    // tau) in FP64 on the FP32 values, each result rounded once when stored
    int tau64 = 0;
};
using MgVars = MgVarsT<double>;

// Stages that take the per-cell kernels only: cylindrical trees and the
// level-set operators (the tiled, paired, tiled-restriction and LDS bottom
// kernels hold the Cartesian lpl and lpld operators).
template <typename R>
inline bool per_cell_only(const DevTree &T, const MgVarsT<R> &V)
{
    return T.cyl_rmin != nullptr || V.lsf != nullptr;
}

// The lower part of a V-cycle -- every level from `hl` down to the lowest
// and back up -- run by one workgroup in one launch (levels with few cells,
// where per-stage launches would cost more than the work).  Levels are
// indexed relative to the lowest one.
constexpr int MAX_BOTTOM_LEVELS = 12;
constexpr int BOTTOM_BLOCK = 512;
// One cell of the lowest level, for the relaxation there (no refinement
// boundaries): its position in the level's array, its box slot and, per face
// in stencil order, where a colour pass after the first reads the
// neighbour: fly >= 0 is a position in the level's array (a cell of this or
// a neighbouring box); fly = -1 (Dirichlet) or -2 (Neumann) is a physical
// boundary, whose ghost is K - phi or K + phi with K = c2 * b (bc_to_gc,
// m_aX_gc.f90:71-132).
struct BaseCell {
    int pos, slot;
    int fly[6];
    double K[6];
};

// Level tables of the bottom kernel (index: level relative to the lowest
// one), in device memory that the kernel only reads (its own argument,
// declared __restrict__, so that the reads are scalar loads).  The bottom
// levels own slots [0, n_slots), boundary records [0, n_bc) and n_par
// parent-list entries (list of level li at par_off[li]); the kernel copies
// these topology tables
// to the first table_bytes of its LDS.  Levels with in_lds set are copied to
// LDS too, after the tables: variable v (phi, rhs, tmp, eps) of level li
// starts at lds_off[v][li] doubles there; the others stay in the HBM pools.
// Levels 0..wave_top run in the first wavefront alone, with wavefront-level
// ordering in place of workgroup barriers (-1: none).  The relaxation of
// level 0 runs from the cell list `base`.
struct BottomLevels {
    LevelArgs lv[MAX_BOTTOM_LEVELS];
    ListArgs par[MAX_BOTTOM_LEVELS];   // parents of each level
    int box_len[MAX_BOTTOM_LEVELS];    // (nc+2)^dim
    int in_lds[MAX_BOTTOM_LEVELS];
    int lds_off[4][MAX_BOTTOM_LEVELS];
    int par_off[MAX_BOTTOM_LEVELS];
    int n_slots, n_bc, n_par, table_bytes;
    int wave_top;
    const BaseCell *base;   // cells of the lowest level: base_n[0] of colour 1, then colour 2
    int base_n[2];
    // 1: every box of the lowest level has 2^dim cells and physical
    // boundaries on all faces (its cells then run in registers, a box per thread)
    int base_fast;
};
struct BottomArgs {
    int hl;                       // highest level of this call (relative index)
    int n_down, n_up, n_base;
    // FP32 stages: the relaxation of the lowest level (the coarsest-grid
    // solve) in FP64 arithmetic on the FP32 values
    int base64 = 0;
    // FP32 stages: update_coarse in FP64 arithmetic (MgVarsT::tau64)
    int tau64 = 0;
    // when not null: s_memrealtime at the start and after every step (the
    // bottom kernel's thread 0), at most trace_cap values
    unsigned long long *trace = nullptr;
    int trace_cap = 0;
};

// The persistent kernel: the lower part of a V-cycle (levels 0..A.hl, data in
// the HBM pools) in one launch of n_rank workgroups of PERSIST_BLOCK threads
// that all run on one XCD and share its L2, ordered by barriers on the
// counters `sync` ([0] arrivals, [1] departures, [2] error flag: 1 a wait
// timed out, 2 the workgroups were not on one XCD; zero between launches;
// [4 + rank] the XCD of each workgroup, plus one).  The grid has 8 * n_rank workgroups; those with
// blockIdx % 8 == xcd take part (workgroups go to the XCDs round-robin).
constexpr int PERSIST_BLOCK = 256;
constexpr int PERSIST_MAX_RANKS = 32;
struct PersistArgs {
    BottomArgs A;
    int set_residual;   // the residual of levels 0..A.hl as the last stage
    int n_rank, xcd;
    unsigned *sync;
    // when not null: s_memrealtime at the start and after every barrier (the
    // first workgroup's view), at most trace_cap values
    unsigned long long *trace = nullptr;
    int trace_cap = 0;
};

}  // namespace afg
