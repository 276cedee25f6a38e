/*
 * afivo_amd.h -- C ABI of the MI355X (gfx950) FAS multigrid library.
 *
 * The library is a drop-in for afivo's multigrid path: the Fortran module
 * m_a2_mg / m_a3_mg of this project (afivo_amd/fortran) keeps the reference's
 * public names and calls these entry points through ISO_C_BINDING.  All
 * arguments are plain C types; every entry point returns 0 on success and a
 * non-zero afg_status otherwise, with the message available from
 * afg_last_error().
 *
 * Conventions follow the reference (src/m_aX_t.f90:153-207):
 *   - box ids are 1-based; a neighbour id > 0 is a box, 0 (a5_no_box) a
 *     refinement boundary and < 0 a physical boundary; a child id 0 means
 *     no child;
 *   - variable indices iv are 1-based, as box%cc(..., iv);
 *   - a box on level lvl has nc_l = nc (lvl >= 1) or nc / 2^(1-lvl) cells per
 *     dimension and stores cc(0:nc_l+1, ...) with one ghost layer, x fastest.
 *
 * Threading: a context is used by one host thread at a time.  Work is queued
 * on the context's own HIP stream; entry points that return data to the host
 * synchronise that stream.
 */
#ifndef AFIVO_AMD_H
#define AFIVO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afg_ctx afg_ctx;

enum afg_status {
    AFG_OK = 0,
    AFG_ERR_ARG = 1,        /* invalid argument */
    AFG_ERR_STATE = 2,      /* call out of order (e.g. no tree set) */
    AFG_ERR_HIP = 3,        /* HIP runtime error */
    AFG_ERR_MEMORY = 4,     /* device memory exhausted / over budget */
    AFG_ERR_UNSUPPORTED = 5 /* operator or configuration not implemented on device */
};

/* Box operator families of m_aX_mg (mg$D_t procedure pointers). */
enum afg_op {
    AFG_OP_LPL = 0,   /* mg$D_box_lpl / gsrb_lpl / corr_lpl / rstr_lpl   (m_aX_mg.f90:708-808) */
    AFG_OP_AUTO = 1,  /* mg$D_auto_op / auto_gsrb / auto_corr (m_aX_mg.f90:562-665): per box by its tag,
                         lpl, lpld (eps) or lpllsf (level set) */
    AFG_OP_LSF = 2    /* mg$D_box_lpllsf / gsrb_lpllsf / corr_lpllsf for every box (m_aX_mg.f90:1020-1253) */
};

/* Restriction of the multigrid (mg%box_rstr). */
enum afg_rstr {
    AFG_RSTR_LPL = 0,   /* mg$D_box_rstr_lpl (m_aX_mg.f90:800-808) */
    AFG_RSTR_AUTO = 1,  /* mg$D_auto_rstr (m_aX_mg.f90:632-647): rstr_lpllsf for children tagged
                           mg_lsf_box, else rstr_lpl.  The reference stops on a child whose tag is
                           unset (:639); the device sets every box's tag before the first stage,
                           which is the state of a program that called mg$D_set_box_tag on each box */
    AFG_RSTR_LSF = 2    /* mg$D_box_rstr_lpllsf for every box (m_aX_mg.f90:1256-1321) */
};

/* Physical boundary types (a5_bc_dirichlet, a5_bc_neumann; m_aX_t.f90:45-48). */
enum { AFG_BC_DIRICHLET = -1, AFG_BC_NEUMANN = -2 };

/* Coordinate systems (a5_xyz, a5_cyl; m_aX_t.f90:50-53). */
enum { AFG_COORD_XYZ = 0, AFG_COORD_CYL = 1 };

/* Refinement-boundary methods of afg_gc_tree / afg_gc_level (the subr_rb
 * argument of a$D_gc_tree, m_aX_gc.f90:28-69). */
enum {
    AFG_RB_MG = 0,          /* mg$D_sides_rb (m_aX_mg.f90:262-350), as afg_fill_gc */
    AFG_RB_PROLONG0 = 1,    /* a$D_gc_prolong0 (m_aX_gc.f90:135-151) */
    AFG_RB_INTERP = 2,      /* a$D_gc_interp (m_aX_gc.f90:155-251) */
    AFG_RB_INTERP_LIM = 3   /* a$D_gc_interp_lim (m_aX_gc.f90:254-358) */
};

/* Arithmetic of the relaxation, transfer and correction stages. */
/* The multigrid options of mg$D_t (m_aX_mg.f90:20-52).  Unset values are -1,
 * filled in by afg_mg_init exactly as mg$D_init_mg does (m_aX_mg.f90:130-154). */
typedef struct afg_mg_opts {
    int32_t i_phi, i_rhs, i_tmp;     /* required */
    int32_t i_eps;                   /* AFG_OP_AUTO only, else -1 */
    int32_t n_cycle_down, n_cycle_up, n_cycle_base;
    int32_t op;                      /* enum afg_op */
    int32_t initialized;             /* set by afg_mg_init */
    /* level set and the boundary value on its zero contour (mg%i_lsf,
     * mg%i_bval; both or neither; -1 or 0: unset), for AFG_OP_AUTO tags,
     * AFG_OP_LSF and the level-set restrictions.  Their ghost cells are the
     * caller's, as for eps.  The level-set operators run in FP64 only. */
    int32_t i_lsf, i_bval;
    int32_t rstr;                    /* enum afg_rstr */
    int32_t reserved[4];
} afg_mg_opts;

/* ------------------------------------------------------------------ */
/* context                                                             */

/* Create a context on the current HIP device for trees of dimension dim
 * (2 or 3) with n_cell cells per box dimension and n_var cell variables.
 * Replaces the device side of a$D_init (m_aX_core.f90:24-131). */
int afg_ctx_create(int dim, int n_cell, int n_var, afg_ctx **out);
int afg_ctx_destroy(afg_ctx *ctx);
/* Message of the last failed call on ctx (or of afg_ctx_create when ctx is NULL). */
const char *afg_last_error(const afg_ctx *ctx);
/* Wait for all work queued on the context's stream. */
int afg_sync(afg_ctx *ctx);

/* ------------------------------------------------------------------ */
/* tree                                                                */

/* Set the tree topology: n_boxes int32 records
 *   [id, lvl, ix(dim), parent, children(2^dim), neighbors(2*dim)]
 * (box$D_t components, m_aX_t.f90:153-177), in any order.  dr_base is the
 * cell size at level 1 (a$D_t%dr_base).  Builds the per-level tables in
 * Morton order (m_morton.f90:33-59), allocates the device pools and resets
 * every physical boundary to Dirichlet zero.  Existing device data is lost. */
int afg_set_tree(afg_ctx *ctx, int n_boxes, const int32_t *records, double dr_base);

/* Batched solves: a forest of independent trees (same dim, n_cell and
 * dr_base) in one context.  As afg_set_tree, plus tree_of[r] in 0..n_trees-1
 * for record r; a box's parent, children and neighbours must be in its own
 * tree.  Every stage then runs over the boxes of a level in all trees in one
 * launch.  afg_fas_fmg and afg_fas_vcycle give every tree exactly the result
 * of solving it alone (each tree's FMG stops at its own highest level and
 * stores its residual there; a V-cycle from highest_lvl runs on each tree
 * from min(highest_lvl, its highest level)).  No counterpart in the
 * reference: afivo solves one tree per call. */
int afg_set_forest(afg_ctx *ctx, int n_boxes, const int32_t *records, double dr_base,
                   const int32_t *tree_of);
/* Replace the tree with its refined / derefined successor, keeping the data on
 * the device: the data side of a$D_adjust_refinement (m_aX_core.f90:452-541)
 * followed by the prolongation of the new boxes that afivo's programs do
 * (e.g. examples/random_refinement_2d.f90:121-144).  records / dr_base /
 * tree_of describe the new tree as for afg_set_forest (tree_of may be NULL
 * for one tree; tree indices must denote the same trees as before).  A box
 * of the new tree that the old tree also had (same tree, level and ix; its
 * id may differ) keeps every value of every variable, ghost cells included.
 * A box the update added gets the n_vars variables `vars` set from its
 * parent, lowest level first, with a$D_prolong0 / a$D_prolong1 /
 * a$D_prolong2 (prolong_order 0 / 1 / 2, m_aX_prolong.f90:39-348; -1: none)
 * on its interior; its other values are zero, so call afg_fill_gc before
 * using its ghost cells.  Boundary conditions set with afg_set_bc carry over
 * to the faces of kept boxes; new boundary faces get Dirichlet zero.  Tile
 * shapes and pass fusion chosen for a level carry over; FP32 copies, graphs
 * and mixed-precision state are reset.  On failure the context has no tree
 * (as after a failed afg_set_forest). */
int afg_update_tree(afg_ctx *ctx, int n_boxes, const int32_t *records, double dr_base,
                    const int32_t *tree_of, int prolong_order, int n_vars, const int32_t *vars);
/* Replace the forest with another one that keeps some of its trees: tree t of
 * the new forest (records and tree_of as for afg_set_forest) is tree
 * old_tree[t] of the current one, or new (old_tree[t] = -1).  The boxes of a
 * kept tree keep every cell and face variable (matched by level and ix
 * within the tree, so a kept tree may also have been refined); the boxes of
 * new trees start from zeros, ready for afg_upload_cc.  Boundary conditions
 * of kept physical faces stay; the others are Dirichlet zero.  For serving
 * loops whose forests take in and release trees between solver steps.  No
 * counterpart in the reference. */
int afg_update_forest(afg_ctx *ctx, int n_boxes, const int32_t *records, double dr_base,
                      const int32_t *tree_of, const int32_t *old_tree);
/* The coordinate system of the trees of the context (the coord argument of
 * a$D_init, m_aX_core.f90:24-131; default AFG_COORD_XYZ).  AFG_COORD_CYL (2D
 * only; x is the radius, r_base the radius at the low side of ix = 1):
 * the auto operators (AFG_OP_AUTO) become mg2_box_clpl / mg2_box_gsrb_clpl
 * for boxes without an eps tag and mg2_box_clpld / mg2_box_gsrb_clpld for eps
 * boxes (m_aX_mg.f90:1323-1443), and every restriction is the
 * radius-weighted one of a$D_restrict_box (m_aX_restrict.f90:90-103).
 * Cylindrical trees run the per-cell kernels (no tiles, fused pass pairs or
 * bottom kernel) and FP64 only.  The setting survives afg_set_tree. */
int afg_set_coord(afg_ctx *ctx, int coord, double r_base);
/* Number of trees and each tree's highest level (highest may be NULL). */
int afg_forest_trees(const afg_ctx *ctx, int *n_trees, int32_t *highest);
/* Per tree: max |variable iv| over its leaf interiors, i.e. max(|min|, |max|)
 * of a$D_tree_max_cc / a$D_tree_min_cc (m_aX_utils.f90:468-571) per tree. */
int afg_forest_max_abs(afg_ctx *ctx, int iv, int n_trees, double *out);

/* ------------------------------------------------------------------ */
/* ghost cells and transfers of any variable                           */

/* a$D_gc_tree (m_aX_gc.f90:28-69) for variable iv: the face ghosts of every
 * box, lowest level first; a same-level neighbour gives sides_from_nb
 * (:381-415), a refinement boundary the method rb (AFG_RB_*), a physical
 * boundary bc_to_gc (:71-132) with the boundary table of afg_set_bc (bc = 0)
 * or with a$D_bc_dirichlet_zero / a$D_bc_neumann_zero on every face
 * (bc = AFG_BC_DIRICHLET / AFG_BC_NEUMANN, :361-379).  afg_gc_level does
 * one level of it (the loop body of a$D_gc_tree, :38-43). */
int afg_gc_tree(afg_ctx *ctx, int iv, int rb, int bc);
int afg_gc_level(afg_ctx *ctx, int iv, int lvl, int rb, int bc);
/* a$D_gc2_box (m_aX_gc.f90:420-456) of variable iv for the boxes ids: the
 * second layer of ghost values beyond each face, with a$D_gc2_prolong1 at
 * refinement boundaries (:498-597) and a$D_bc2_neumann_zero at physical
 * boundaries (:600-634), packed per box as gc_data(nc_l[, nc_l], 2*dim) in
 * Fortran order (first tangential index fastest, then the face).  It reads
 * the neighbours' cells and, at refinement boundaries, the parent's
 * neighbour with its ghost cells: fill those first (afg_gc_tree). */
int afg_gc2_boxes(afg_ctx *ctx, int iv, int n_ids, const int32_t *ids, double *packed);
/* a$D_restrict_tree (m_aX_restrict.f90:47-57): children of every parent,
 * highest level first, variable iv restricted into iv_to of the parents. */
int afg_restrict_tree(afg_ctx *ctx, int iv, int iv_to);
/* a$D_prolong0 / a$D_prolong1 / a$D_prolong2 (order 0 / 1 / 2,
 * m_aX_prolong.f90:39-348) of variable iv from the parents of level lvl into
 * iv_to of every box of level lvl, i.e. a$D_prolong<order>_from for every
 * parent of level lvl - 1; add = 1 adds to the old values. */
int afg_prolong_level(afg_ctx *ctx, int iv, int iv_to, int lvl, int order, int add);
/* Whole-array arithmetic over every box (ghost cells included):
 * a$D_box_add_cc (iv_to += iv_from), a$D_box_sub_cc (iv_to -= iv_from),
 * a$D_box_times_cc (iv *= a), a$D_box_lincomb_cc (iv_b = a*iv_a + b*iv_b),
 * m_aX_utils.f90:273-318, applied to the whole tree. */
int afg_add_cc(afg_ctx *ctx, int iv_from, int iv_to);
int afg_sub_cc(afg_ctx *ctx, int iv_from, int iv_to);
int afg_times_cc(afg_ctx *ctx, double a, int iv);
int afg_lincomb_cc(afg_ctx *ctx, double a, int iv_a, double b, int iv_b);

/* Face variables ifc = 1..AFG_MAX_FC (box%fx, fy (, fz), m_aX_core.f90:
 * 359-369), allocated and zeroed on first use.  Packed data of a box: fx, fy
 * (, fz), each (nc_l+1) * nc_l^(dim-1) values with extent nc_l+1 along its
 * own direction, Fortran order; boxes in the order of `ids`. */
#define AFG_MAX_FC 16
int afg_upload_fc(afg_ctx *ctx, int ifc, int n_ids, const int32_t *ids, const double *packed);
int afg_download_fc(afg_ctx *ctx, int ifc, int n_ids, const int32_t *ids, double *packed);
/* a$D_consistent_fluxes (m_aX_core.f90:811-943, Cartesian) over the n_f
 * face variables f_ixs: on every face a parent shares with a neighbour that
 * has no children, the neighbour's fluxes become the mean of the parent's
 * children's fluxes on that face; lowest level first. */
int afg_consistent_fluxes(afg_ctx *ctx, int n_f, const int32_t *f_ixs);

/* Boundary condition of box `id` in direction nb (1-based, a$D_neighb_*):
 * bc_type AFG_BC_DIRICHLET or AFG_BC_NEUMANN and the boundary values b written
 * by the user's sides_bc (nc_l^(dim-1) values, first tangential index
 * fastest), or NULL for b = 0.  The ghost value is c2*b + c1*interior as in
 * bc_to_gc (m_aX_gc.f90:71-132). */
int afg_set_bc(afg_ctx *ctx, int id, int nb, int bc_type, const double *b);

/* Number of doubles of one variable packed over the given ids (each box its
 * whole cc(0:nc_l+1, ..., iv) array, Fortran order). */
int afg_packed_size(const afg_ctx *ctx, int n_ids, const int32_t *ids, int64_t *n_values);
/* Copy variable iv of boxes ids[0..n_ids) from host (packed) to the device,
 * and back.  The host side of a copy runs on up to 16 threads and overlaps
 * the PCIe transfer; an upload returns once the values are staged (the
 * packed array may be reused), a download once they are in `packed`. */
int afg_upload_cc(afg_ctx *ctx, int iv, int n_ids, const int32_t *ids, const double *packed);
int afg_download_cc(afg_ctx *ctx, int iv, int n_ids, const int32_t *ids, double *packed);
/* The same from and to separate host arrays, one per box: boxes[k] is box
 * ids[k]'s whole cc(0:nc_l+1, ..., iv) array of the variable, x fastest -- the
 * box%cc(..., iv) slab of an afivo box, so that a Fortran caller passes
 * c_loc(box%cc(0, 0, iv)) and packs nothing (the drop-in's transfers of box%cc,
 * m_aX_mg.f90 mg$D_fas_fmg / mg$D_fas_vcycle).  Each id at most once. */
int afg_upload_boxes(afg_ctx *ctx, int iv, int n_ids, const int32_t *ids, const double *const *boxes);
int afg_download_boxes(afg_ctx *ctx, int iv, int n_ids, const int32_t *ids, double *const *boxes);
/* Whole-array device copy cc(..., iv_to) = cc(..., iv_from) on every box, and
 * cc(..., iv) = value on every box (a$D_box_copy_cc / a$D_box_clear_cc). */
int afg_copy_cc(afg_ctx *ctx, int iv_from, int iv_to);
int afg_set_cc(afg_ctx *ctx, int iv, double value);

/* ------------------------------------------------------------------ */
/* multigrid (m_aX_mg)                                                 */

/* mg$D_init_mg (m_aX_mg.f90:130-154): check and default the options. */
int afg_mg_init(afg_mg_opts *opts);
/* mg$D_fas_fmg (m_aX_mg.f90:164-203). */
int afg_fas_fmg(afg_ctx *ctx, const afg_mg_opts *opts, int set_residual, int have_guess);
/* mg$D_fas_vcycle (m_aX_mg.f90:208-255). */
int afg_fas_vcycle(afg_ctx *ctx, const afg_mg_opts *opts, int highest_lvl, int set_residual);
/* Highest level present in the tree (a$D_t%highest_lvl) and the lowest
 * (lbound(tree%lvls)). */
int afg_levels(const afg_ctx *ctx, int *min_lvl, int *highest_lvl);
/* The ids of the boxes of level lvl in the order the device holds them (the
 * level's slot range), into ids[0..n) -- tree%lvls(lvl)%ids (m_aX_t.f90:
 * 145-149) sorted as a$D_tidy_up sorts it (m_aX_core.f90:272-279, key
 * morton_from_ix$D(ix - 1), m_morton.f90:33-59); in a forest grouped by
 * tree first (trees of higher highest level first, then tree index).  n
 * must be at least the level's box count; *n_out receives it. */
int afg_level_ids(const afg_ctx *ctx, int lvl, int n, int32_t *ids, int *n_out);

/* a$D_tree_max_cc / a$D_tree_min_cc (m_aX_utils.f90:468-571): extreme of
 * cc(..., iv) over leaf interiors (or all boxes if include_parents). */
int afg_tree_max_cc(afg_ctx *ctx, int iv, int include_parents, double *out);
int afg_tree_min_cc(afg_ctx *ctx, int iv, int include_parents, double *out);

/* Largest number of per-block partial results the allocation-free reduction
 * below holds (doubles of its scratch array, 128 KiB per device). */
#define AFG_NOALLOC_MAX_PARTIALS 16384
/* Allocation-free variant of afg_tree_max_cc (want_max = 1) and
 * afg_tree_min_cc (want_max = 0): the same value, bit for bit, written to
 * d_out[0] in device memory, stream-ordered on the context's stream (no
 * synchronisation, no host copy).  It calls no hipMalloc or hipFree and
 * nothing that allocates: the per-block partial results go to a scratch
 * array of AFG_NOALLOC_MAX_PARTIALS doubles that the library holds on each
 * device until the process ends.  Trees that need more partials are
 * rejected (AFG_ERR_UNSUPPORTED), and so is a variable that has no device
 * data yet (AFG_ERR_STATE: creating it would allocate).  Every call on a
 * device shares that device's scratch array, which keeps the last call's
 * partials: calls that overlap on one device (several contexts, streams or
 * host threads at once) race; calls in one stream do not.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). */
int afg_tree_extreme_noalloc(afg_ctx *ctx, int iv, int include_parents, int want_max, double *d_out);

/* Single stages over one level (the building blocks of the cycles above,
 * exported for stage-level checks and for callers that drive their own
 * cycles). */
/* fill_gc_phi (m_aX_mg.f90:352-365) = a$D_gc_box over a level with
 * mg$D_sides_rb and the boundary table (m_aX_gc.f90:50-69). */
int afg_fill_gc(afg_ctx *ctx, int iv, int lvl);
/* gsrb_boxes (m_aX_mg.f90:397-420). */
int afg_gsrb_level(afg_ctx *ctx, const afg_mg_opts *opts, int lvl, int n_cycle);
/* residual_box over a level (m_aX_mg.f90:545-559). */
int afg_residual_level(afg_ctx *ctx, const afg_mg_opts *opts, int lvl);
/* a$D_restrict_box of every box of lvl into its parent (m_aX_restrict.f90:67-132). */
int afg_restrict_level(afg_ctx *ctx, int iv, int lvl);
/* a$D_prolong2 from the parent into every box of lvl (m_aX_prolong.f90:243-348). */
int afg_prolong2_level(afg_ctx *ctx, int iv, int iv_to, int lvl, int add);
/* correct_children over the parents of lvl-1 (m_aX_mg.f90:368-395). */
int afg_correct_children(afg_ctx *ctx, const afg_mg_opts *opts, int lvl);
/* update_coarse (m_aX_mg.f90:424-480). */
int afg_update_coarse(afg_ctx *ctx, const afg_mg_opts *opts, int lvl);

/* ------------------------------------------------------------------ */
/* mixed precision                                                     */

/* FP32 copies of the cell variables, for the FP32 stages: variable iv32 of the
 * FP32 pool = float(variable iv), and back (added to variable iv if add). */
int afg_fp32_from_cc(afg_ctx *ctx, int iv, int iv32);
int afg_fp32_to_cc(afg_ctx *ctx, int iv32, int iv, int add);
/* mg$D_fas_fmg / mg$D_fas_vcycle on the FP32 pool: the variables of opts are
 * FP32 variables; the same stages in FP32 arithmetic (tags of the auto
 * operators come from the FP64 eps). */
int afg_fas_fmg_fp32(afg_ctx *ctx, const afg_mg_opts *opts, int set_residual, int have_guess);
int afg_fas_vcycle_fp32(afg_ctx *ctx, const afg_mg_opts *opts, int highest_lvl, int set_residual);
/* Solve A phi = rhs (the options' operator, phi, rhs, tmp, eps in FP64) by
 * FP64 defect correction, in steps (no counterpart in the reference):
 * afg_mixed_begin sets phi = 0 unless have_guess and stores the FP64
 * residual r = rhs - A phi in tmp on every level, phi of every parent first
 * set to the restriction of its children's (the FAS fixed point).  Each
 * afg_mixed_correct computes a correction e from A e = r by FP32 FAS
 * multigrid in the FP32 copies of the same variables (`cycles` V-cycles from
 * e = 0; the first correction after afg_mixed_begin starts with an FMG),
 * adds it to phi and stores the new FP64 residual.  The caller checks the
 * residual (afg_forest_max_abs of tmp).  Physical boundaries must be
 * homogeneous (b = 0: the correction solves with the same ghost rules).
 * Forests: every tree is corrected together, each with its own stages. */
int afg_mixed_begin(afg_ctx *ctx, const afg_mg_opts *opts, int have_guess);
int afg_mixed_correct(afg_ctx *ctx, const afg_mg_opts *opts, int cycles);
/* Whether the first correction after afg_mixed_begin starts with an FMG
 * (default 1).  With 0 every correction is `cycles` V-cycles from e = 0, so
 * the corrections a tree receives do not depend on when afg_mixed_begin was
 * last called: afg_mixed_begin with have_guess after afg_update_forest
 * leaves the kept trees' phi and residual bit for bit as they were. */
int afg_mixed_first_fmg(afg_ctx *ctx, int enabled);
/* The FP32 stages' coarsest-grid solve (the n_cycle_base sweeps on the
 * lowest level) in FP64 arithmetic on the FP32 values, kept in registers
 * across the sweeps (default 0: FP32).  It applies where the lowest level's
 * boxes have 2^dim cells and physical boundaries on every face (one or more
 * base boxes of the coarsened tree, as afivo builds them with coarsen_to = 2);
 * elsewhere the coarsest solve stays FP32. */
int afg_mixed_fp64_base(afg_ctx *ctx, int enabled);
/* The FP32 stages' update_coarse (m_aX_mg.f90:424-480: the fine residual,
 * its restriction and that of phi, and rhs = A phi + tmp on the parents,
 * which together form the FAS tau) in FP64 arithmetic on the FP32 values,
 * each stored result rounded once to FP32 (default 0: FP32 arithmetic).
 * Cartesian operators only, as the FP32 stages. */
int afg_mixed_fp64_tau(afg_ctx *ctx, int enabled);
/* The whole solve: afg_mixed_begin, then corrections until every tree has
 * max|rhs - A phi| <= tol * max|rhs| over its leaf interiors, or max_iter
 * corrections.  Returns the number of corrections and the largest relative
 * residual over the trees; tmp holds the FP64 residual. */
int afg_mixed_solve(afg_ctx *ctx, const afg_mg_opts *opts, int have_guess, double tol, int max_iter,
                    int cycles, int *iterations, double *rel_residual);

/* ------------------------------------------------------------------ */
/* execution controls                                                  */

/* The lowest levels of the tree, while each holds at most `cells` cells
 * (default 2048) and all of them fit in one workgroup's LDS, run their part
 * of every V-cycle in one single-workgroup launch; 0 turns this off.  Within
 * it, the lowest levels holding at most `wave_cells` cells each (default 0)
 * run in a single wavefront.  Sets the limit of both precisions.  Results
 * are identical either way. */
int afg_set_bottom_cells(afg_ctx *ctx, int64_t cells, int64_t wave_cells);
/* The cell limit in use for the FP64 (fp32 = 0) or FP32 stages. */
int afg_get_bottom_cells(const afg_ctx *ctx, int fp32, int64_t *cells);
/* Relaxation blocks of two or more single colour passes on 3D levels with
 * 8^3 or 16^3 boxes keep the level's phi by colour in a working copy between
 * their passes (the first pass reads the box arrays, the last writes them);
 * enabled = 0 runs every pass on the box arrays.  Default on.  Results are
 * identical either way.  No counterpart in the reference (an execution
 * control of gsrb_boxes, m_aX_mg.f90:397-420). */
int afg_set_split_passes(afg_ctx *ctx, int enabled);
/* The persistent kernel: with wgs in 1..32, the lowest levels of the tree,
 * while each holds at most `cells` cells, run their part of every V-cycle
 * (mg$D_fas_vcycle below and at the highest such level, m_aX_mg.f90:218-241,
 * with the residual when the V-cycle's highest level is among them) in one
 * launch of wgs workgroups that run on one XCD and meet at barriers between
 * the stages, in place of the bottom kernel and the per-level launches of
 * those levels; wgs = 0 (default) turns it off.  Results are identical
 * either way.  A barrier wait is bounded: should one time out, the next
 * synchronising call fails with AFG_ERR_HIP and the kernel is turned off.
 * No counterpart in the reference (an execution control of the path). */
int afg_set_persist(afg_ctx *ctx, int wgs, int64_t cells);
/* The workgroups (0: off), the cell limit, and the highest level the
 * persistent kernel takes for the FP64 stages (-1000: none). */
int afg_get_persist(const afg_ctx *ctx, int *wgs, int64_t *cells, int *top_lvl);
/* Stage trace of the persistent kernel and of the bottom kernel (a
 * diagnostic): returns in ticks[0..n) the s_memrealtime values (100 MHz) that
 * the first workgroup of the last such launch took at its start, after each
 * barrier between its steps and at its end (the bottom kernel) since the previous call,
 * when ticks and n are given (ticks holds the capacity set by the previous
 * call); then records at most `cap` values per launch from now on (cap = 0:
 * none, the default). */
int afg_persist_trace(afg_ctx *ctx, int cap, int64_t *ticks, int *n);
/* Shape of the tiled colour pass on level lvl, for the FP64 (fp32 = 0) or
 * FP32 stages: 0 (default), 1 or 2 for nc 8 and 16 (workgroup size and boxes
 * per workgroup), 0 otherwise.  Results are identical for every shape. */
int afg_set_tile(afg_ctx *ctx, int lvl, int fp32, int shape);
int afg_get_tile(const afg_ctx *ctx, int lvl, int fp32, int *shape);
/* Autotuner, for the FP64 (fp32 = 0) or FP32 stages: on every level with
 * several tile shapes, time `reps` times two colour passes with each shape,
 * as two single passes and as one fused pair (on a copy of phi), and keep
 * the fastest shape and mode (the current ones unless another is more than
 * 3% faster); then time `reps` V-cycles with each bottom cell limit of 512,
 * 2048, 8192 and 32768 and keep the fastest.  The tree's data is unchanged
 * (saved and restored).  If ms is given it receives, per level from the
 * lowest, the milliseconds per two passes of shapes 0..2 as single passes,
 * then as pairs (6 values per level), then the milliseconds per V-cycle of
 * the 4 limits (-1: not timed).  No counterpart in the reference. */
int afg_autotune(afg_ctx *ctx, const afg_mg_opts *opts, int fp32, int reps, double *ms);
/* Run the colour passes of the relaxation blocks (nc 8 and 16) of every
 * level as fused pairs, two passes per launch with phi going to a second
 * pool and back, or one pass per launch (default: pairs in 2D, single
 * passes in 3D; afg_autotune chooses per level).  Results are identical
 * either way. */
int afg_set_pass_pairs(afg_ctx *ctx, int enabled);
int afg_get_pass_pairs(const afg_ctx *ctx, int lvl, int fp32, int *enabled);
/* The same for one level and precision (what afg_autotune chooses per level). */
int afg_set_level_pass_pairs(afg_ctx *ctx, int lvl, int fp32, int enabled);
/* Replay whole V-cycles from HIP graphs (default on). */
int afg_set_graphs(afg_ctx *ctx, int enabled);
/* Per-stage device timing: while enabled, each launch of the GSRB colour-pass
 * kernel on level `lvl` (every level when lvl is below the lowest) is
 * bracketed with HIP events on the context stream; afg_timing reads the
 * number of launches, their summed duration (ms) and their summed
 * algorithmic bytes.  Results are identical with timing on or off. */
int afg_set_timing(afg_ctx *ctx, int enabled, int lvl);
int afg_timing(afg_ctx *ctx, int64_t *n_launches, double *ms, double *bytes);
/* Number of cell-updates (cells relaxed by one full red+black sweep) one
 * V-cycle with these options performs, summed over levels. */
int afg_vcycle_cell_updates(const afg_ctx *ctx, const afg_mg_opts *opts, int highest_lvl, double *out);
/* Device bytes held by the context (pools and tables). */
int afg_device_bytes(const afg_ctx *ctx, int64_t *out);

/* ------------------------------------------------------------------ */
/* Per-box operators on the host (CPU): the public box procedures of the
 * reference module for one box in host memory, for callers that apply them
 * to single boxes (the drop-in module's mg$D_box_* procedures).  They run the
 * per-cell bodies of the device stages, so each gives the device stage's
 * value for that box.  Box arrays hold cc(0:nc+1, 0:nc+1[, 0:nc+1]) of one
 * variable with x fastest; dr is box%dr; ir0 = (box%ix(1) - 1) * nc (the
 * radial cell offset, cylindrical kinds only).  No context, no device. */
#define AFG_BOX_LPL 0   /* mg$D_box_lpl / gsrb_lpl / corr_lpl / rstr_lpl */
#define AFG_BOX_LPLD 1  /* mg$D_box_lpld / gsrb_lpld / corr_lpld */
#define AFG_BOX_LSF 2   /* mg$D_box_lpllsf / gsrb_lpllsf / corr_lpllsf / rstr_lpllsf */
#define AFG_BOX_CLPL 3  /* mg2_box_clpl / gsrb_clpl; restriction with the radial weights */
#define AFG_BOX_CLPLD 4 /* mg2_box_clpld / gsrb_clpld */
/* out(1:nc, ...) = A phi: mg$D_box_lpl (m_aX_mg.f90:763-797), _lpld
 * (:866-916), _lpllsf (:1183-1253), mg2_box_clpl / _clpld (:1354-1407).
 * eps for the lpld kinds, lsf and bval for AFG_BOX_LSF (else may be NULL). */
int afg_box_op(int dim, int nc, double dr, int kind, int ir0, const double *phi, const double *eps,
               const double *lsf, const double *bval, double *out);
/* One colour of Gauss-Seidel red-black on phi: mg$D_box_gsrb_lpl
 * (m_aX_mg.f90:718-760), _lpld (:810-863), _lpllsf (:1103-1180),
 * mg2_box_gsrb_clpl (:1325-1351) / _clpld (:1411-1443); the parity of redblack selects the
 * cells, as redblack_cntr does. */
int afg_box_gsrb(int dim, int nc, double dr, int kind, int ir0, int redblack, double *phi, const double *rhs,
                 const double *eps, const double *lsf, const double *bval);
/* Correction of a child from its parent's tmp (the parent's correction):
 * mg$D_box_corr_lpl (m_aX_mg.f90:708-715), _lpld (:918-985), _lpllsf
 * (:1020-1100), adding into the child's phi.  offset is
 * a$D_get_child_offset(box_c) (m_aX_t.f90); pnc the parent's cell count;
 * p_eps for AFG_BOX_LPLD, p_lsf and c_lsf for AFG_BOX_LSF. */
int afg_box_corr(int dim, int nc, int pnc, int kind, const int32_t *offset, const double *p_corr,
                 const double *p_eps, const double *p_lsf, double *c_phi, const double *c_lsf);
/* Restriction of a child into its parent: mg$D_box_rstr_lpl =
 * a$D_restrict_box (m_aX_restrict.f90:67-132; AFG_BOX_CLPL: the radial
 * weights of a cylindrical parent, p_rmin = box_p%r_min(1), c_dr = box_c%dr)
 * and mg$D_box_rstr_lpllsf (m_aX_mg.f90:1256-1321; c_lsf, p_lsf). */
int afg_box_rstr(int dim, int nc, int pnc, int kind, const int32_t *offset, const double *c_var,
                 const double *c_lsf, const double *p_lsf, double p_rmin, double c_dr, double *p_var);
/* mg$D_sides_rb (m_aX_mg.f90:262-350) of face nb (1-based, afivo's
 * neighbour order) of a child from its parent (ghost cells included). */
int afg_box_sides_rb(int dim, int nc, int pnc, int nb, const int32_t *offset, const double *p_var, double *c_var);

#ifdef __cplusplus
}
#endif
#endif
