/*
 * afivo_oracle.h -- CPU restatement of afivo's FAS multigrid path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against; the product library never links or
 * calls it.  Every routine cites the reference routine it restates.
 *
 * Tree input is the same flat record format the product C ABI takes:
 * one int32 record per box,
 *     [id, lvl, ix(dim), parent, children(2^dim), neighbors(2*dim)]
 * with afivo's conventions (1-based ids; a neighbour id > 0 is a box, 0 is a
 * refinement boundary, < 0 a physical boundary; child 0 means no child).
 * Cell data of variable iv is packed in record order, each box contributing
 * its whole cc(0:nc_l+1, ..., iv) array in Fortran (x fastest) order, where
 * nc_l = nc for lvl >= 1 and nc / 2^(1-lvl) below.
 */
#ifndef AFIVO_ORACLE_H
#define AFIVO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tree orc_tree;

typedef struct orc_mg_opts {
    int32_t i_phi, i_rhs, i_tmp, i_eps;     /* 1-based variable indices */
    int32_t n_cycle_down, n_cycle_up, n_cycle_base;
    int32_t op;                             /* 0: lpl set, 1: auto set, 2: lpllsf set */
    /* level set and boundary value (mg%i_lsf, mg%i_bval; -1 unset) and the
     * restriction: 0 mg$D_box_rstr_lpl, 1 mg$D_auto_rstr (by the child's
     * tag), 2 mg$D_box_rstr_lpllsf */
    int32_t i_lsf, i_bval, rstr;
} orc_mg_opts;

enum { ORC_BC_DIRICHLET = -1, ORC_BC_NEUMANN = -2 };
#define ORC_MAX_FC 16   /* face variables */
/* refinement-boundary ghost methods of orc_gc_tree */
enum { ORC_RB_MG = 0, ORC_RB_PROLONG0 = 1, ORC_RB_INTERP = 2, ORC_RB_INTERP_LIM = 3 };

orc_tree *orc_create(int dim, int nc, int n_var, double dr_base, int n_boxes,
                     const int32_t *recs);
void orc_destroy(orc_tree *t);
int orc_n_values(const orc_tree *t);          /* doubles per packed variable */
int orc_set_cc(orc_tree *t, int iv, const double *packed);
int orc_get_cc(const orc_tree *t, int iv, double *packed);
/* Physical boundary of box `id` in direction nb (1-based): bc_type and the
 * boundary values b (nc_l^(dim-1), x fastest), or NULL for b = 0. */
int orc_set_bc(orc_tree *t, int id, int nb, int bc_type, const double *b);

/* Cylindrical coordinates (2D only; x is the radius, a5_cyl): the auto
 * operators become mg2_box_clpl / clpld and the restriction is weighted by
 * radius; r_base is the tree's r_min(1) at ix = 1. */
int orc_set_coord(orc_tree *t, int cyl, double r_base);

int orc_fas_fmg(orc_tree *t, const orc_mg_opts *o, int set_residual, int have_guess);
/* every box's tag (mg$D_set_box_tag on each box), as mg$D_auto_rstr needs */
int orc_set_tags(orc_tree *t, const orc_mg_opts *o);
/* One box procedure on one box (per-box checks): what 0 op into i_out = arg,
 * 1 gsrb with redblack_cntr = arg, 2 corr of child id, 3 rstr of variable
 * arg from child id, 4 sides_rb of face arg; kind 0 lpl, 1 lpld, 2 lpllsf,
 * 3 clpl, 4 clpld (cylindrical 2D trees). */
int orc_box_call(orc_tree *t, const orc_mg_opts *o, int what, int kind, int id, int arg);
int orc_fas_vcycle(orc_tree *t, const orc_mg_opts *o, int highest_lvl, int set_residual);
double orc_tree_max_cc(const orc_tree *t, int iv, int include_parents);
double orc_tree_min_cc(const orc_tree *t, int iv, int include_parents);

/* Single stages over one level, for stage-level parity checks. */
int orc_fill_gc_level(orc_tree *t, int iv, int lvl);
int orc_gsrb_level(orc_tree *t, const orc_mg_opts *o, int lvl, int n_cycle);
int orc_residual_level(orc_tree *t, const orc_mg_opts *o, int lvl);
int orc_restrict_level(orc_tree *t, int iv, int lvl);          /* lvl -> lvl-1 */
int orc_prolong2_level(orc_tree *t, int iv, int iv_to, int lvl, int add); /* lvl-1 -> lvl */
int orc_correct_children_level(orc_tree *t, const orc_mg_opts *o, int lvl);

/* Ghost cells and transfers outside the multigrid: a$D_gc_tree with a choice
 * of refinement-boundary method and boundary condition (bc 0: the boundary
 * table; ORC_BC_*: that condition with b = 0 everywhere), a$D_restrict_tree,
 * and a$D_prolong0/1/2 into every box of level lvl (order 0/1/2). */
int orc_gc_tree(orc_tree *t, int iv, int rb, int bc);
/* a$D_gc2_box with a$D_gc2_prolong1 and a$D_bc2_neumann_zero of variable iv
 * on every box, record order: per box gc_data(nc_l[, nc_l], 2*dim), Fortran
 * order (the caller fills the first ghost layer first). */
int orc_gc2(orc_tree *t, int iv, double *out);
int orc_restrict_tree(orc_tree *t, int iv, int iv_to);
int orc_prolong_level(orc_tree *t, int iv, int iv_to, int lvl, int order, int add);
/* Face variables 1..ORC_MAX_FC: per box fx, fy (, fz), each (nc_l+1) *
 * nc_l^(dim-1) values in Fortran order, boxes in record order; and
 * a$D_consistent_fluxes over the face variables f_ixs. */
int orc_n_face_values(const orc_tree *t);
int orc_set_fc(orc_tree *t, int ifc, const double *packed);
int orc_get_fc(const orc_tree *t, int ifc, double *packed);
int orc_consistent_fluxes(orc_tree *t, int n_f, const int *f_ixs);
int orc_update_coarse_level(orc_tree *t, const orc_mg_opts *o, int lvl);

#ifdef __cplusplus
}
#endif
#endif
