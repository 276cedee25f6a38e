/*
 * afivo_oracle.c -- CPU restatement of afivo's FAS multigrid hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the HIP path is compared
 * against (tests, the smoke check and the CPU baseline of bench.py).  The
 * product library never links or calls it.
 *
 * Each function restates one reference routine, cited as file:line in the
 * reference's dimension-generic sources (src/m_aX_*.f90).  Floating-point
 * expressions keep the reference's left-to-right evaluation order and are
 * compiled without contraction (-ffp-contract=off), so that results can be
 * compared bit for bit with the reference built by flang -O2.
 *
 * Box arrays use the Fortran layout of box%cc(0:n+1, 0:n+1[, 0:n+1], iv):
 * x fastest.
 */
#include "afivo_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_tree {
    int dim, nc, n_var, nch, nnb;
    int min_lvl, highest_lvl, n_boxes, max_id;
    double dr_base;
    /* per id, 0..max_id */
    int *lvl, *ix, *parent, *children, *nbr, *ncell, *tag;
    double *dr;
    size_t *off;           /* offset (doubles) of the box array in data[iv] */
    size_t n_values;
    double **data;         /* data[iv-1] */
    int *rec_ids;          /* ids in record order */
    int *lvl_start;        /* record-order ids of level l: [lvl_start[l-min], lvl_start[l-min+1]) */
    int *bc_type;          /* per (id, nb) */
    double **bc_val;       /* per (id, nb), NULL means b = 0 */
    /* face variables: fx, fy (, fz) of each box, packed in record order at
     * foff[id]; allocated on first use */
    size_t *foff;
    size_t n_face;
    double *fc[ORC_MAX_FC];
    /* cylindrical coordinates (2D, x the radius): box%r_min(1) per id */
    int cyl;
    double *rmin;
    /* set when mg$D_set_box_tag finds both a level-set and an eps tag */
    int tag_err;
};

#define TAG_UNSET 0
#define TAG_NORMAL 1   /* mg_normal_box, m_aX_mg.f90:14 */
#define TAG_LSF 2      /* mg_lsf_box, m_aX_mg.f90:15 */
#define TAG_CEPS 3     /* mg_ceps_box, m_aX_mg.f90:16 */
#define TAG_VEPS 4     /* mg_veps_box, m_aX_mg.f90:17 */

static int box_len(const orc_tree *t, int n)
{
    int m = n + 2;
    return t->dim == 2 ? m * m : m * m * m;
}

static double *cc(const orc_tree *t, int id, int iv)
{
    return t->data[iv - 1] + t->off[id];
}

/* Linear index of cell (i, j, k) in a box with n cells per dimension. */
static inline int ix3(int n, int i, int j, int k)
{
    return i + (n + 2) * (j + (n + 2) * k);
}

/* --------------------------------------------------------------------- */
/* tree set-up                                                            */

orc_tree *orc_create(int dim, int nc, int n_var, double dr_base, int n_boxes,
                     const int32_t *recs)
{
    if ((dim != 2 && dim != 3) || nc < 2 || (nc & 1) || n_var < 1 || n_boxes < 1)
        return NULL;
    orc_tree *t = calloc(1, sizeof(*t));
    t->dim = dim;
    t->nc = nc;
    t->n_var = n_var;
    t->nch = 1 << dim;
    t->nnb = 2 * dim;
    t->dr_base = dr_base;
    t->n_boxes = n_boxes;
    int reclen = 3 + dim + t->nch + t->nnb;
    int max_id = 0, min_lvl = 1 << 30, hi_lvl = -(1 << 30);
    for (int r = 0; r < n_boxes; r++) {
        const int32_t *q = recs + (size_t)r * reclen;
        if (q[0] > max_id) max_id = q[0];
        if (q[1] < min_lvl) min_lvl = q[1];
        if (q[1] > hi_lvl) hi_lvl = q[1];
    }
    t->max_id = max_id;
    t->min_lvl = min_lvl;
    t->highest_lvl = hi_lvl;
    size_t n = (size_t)max_id + 1;
    t->lvl = calloc(n, sizeof(int));
    t->ix = calloc(n * 3, sizeof(int));
    t->parent = calloc(n, sizeof(int));
    t->children = calloc(n * 8, sizeof(int));
    t->nbr = calloc(n * 6, sizeof(int));
    t->ncell = calloc(n, sizeof(int));
    t->tag = calloc(n, sizeof(int));
    t->dr = calloc(n, sizeof(double));
    t->off = calloc(n, sizeof(size_t));
    t->rec_ids = calloc((size_t)n_boxes, sizeof(int));
    t->bc_type = calloc(n * 6, sizeof(int));
    t->bc_val = calloc(n * 6, sizeof(double *));
    size_t off = 0;
    for (int r = 0; r < n_boxes; r++) {
        const int32_t *q = recs + (size_t)r * reclen;
        int id = q[0];
        t->rec_ids[r] = id;
        t->lvl[id] = q[1];
        for (int d = 0; d < dim; d++) t->ix[id * 3 + d] = q[2 + d];
        t->parent[id] = q[2 + dim];
        for (int c = 0; c < t->nch; c++) t->children[id * 8 + c] = q[3 + dim + c];
        for (int f = 0; f < t->nnb; f++) {
            t->nbr[id * 6 + f] = q[3 + dim + t->nch + f];
            t->bc_type[id * 6 + f] = ORC_BC_DIRICHLET;   /* a$D_bc_dirichlet_zero */
        }
        int l = q[1];
        /* a$D_n_cell, m_aX_utils.f90:676-686; box%dr, m_aX_core.f90:196,764 */
        t->ncell[id] = l >= 1 ? nc : nc / (1 << (1 - l));
        t->dr[id] = dr_base * pow(0.5, (double)(l - 1));
        t->off[id] = off;
        off += (size_t)box_len(t, t->ncell[id]);
    }
    t->n_values = off;
    t->foff = calloc(n, sizeof(size_t));
    size_t fo = 0;
    for (int r2 = 0; r2 < n_boxes; r2++) {
        int id = t->rec_ids[r2], m = t->ncell[id];
        t->foff[id] = fo;
        fo += (size_t)dim * (m + 1) * (dim == 3 ? m * m : m);
    }
    t->n_face = fo;
    t->data = calloc((size_t)n_var, sizeof(double *));
    for (int v = 0; v < n_var; v++) t->data[v] = calloc(off, sizeof(double));
    int nl = hi_lvl - min_lvl + 1;
    t->lvl_start = calloc((size_t)nl + 1, sizeof(int));
    /* records are required in level order */
    int r = 0;
    for (int l = min_lvl; l <= hi_lvl; l++) {
        t->lvl_start[l - min_lvl] = r;
        while (r < n_boxes && t->lvl[t->rec_ids[r]] == l) r++;
    }
    t->lvl_start[nl] = r;
    if (r != n_boxes) {
        orc_destroy(t);
        return NULL;
    }
    return t;
}

void orc_destroy(orc_tree *t)
{
    if (!t) return;
    if (t->data)
        for (int v = 0; v < t->n_var; v++) free(t->data[v]);
    if (t->bc_val)
        for (size_t i = 0; i < ((size_t)t->max_id + 1) * 6; i++) free(t->bc_val[i]);
    free(t->data); free(t->bc_val); free(t->bc_type);
    free(t->lvl); free(t->ix); free(t->parent); free(t->children); free(t->nbr);
    free(t->ncell); free(t->tag); free(t->dr); free(t->off); free(t->rec_ids);
    free(t->lvl_start);
    free(t->foff);
    free(t->rmin);
    for (int v = 0; v < ORC_MAX_FC; v++) free(t->fc[v]);
    free(t);
}

int orc_n_values(const orc_tree *t) { return (int)t->n_values; }

int orc_set_cc(orc_tree *t, int iv, const double *packed)
{
    if (iv < 1 || iv > t->n_var) return -1;
    memcpy(t->data[iv - 1], packed, t->n_values * sizeof(double));
    return 0;
}

int orc_get_cc(const orc_tree *t, int iv, double *packed)
{
    if (iv < 1 || iv > t->n_var) return -1;
    memcpy(packed, t->data[iv - 1], t->n_values * sizeof(double));
    return 0;
}


int orc_set_bc(orc_tree *t, int id, int nb, int bc_type, const double *b)
{
    if (id < 1 || id > t->max_id || nb < 1 || nb > t->nnb) return -1;
    if (bc_type != ORC_BC_DIRICHLET && bc_type != ORC_BC_NEUMANN) return -1;
    size_t k = (size_t)id * 6 + (nb - 1);
    t->bc_type[k] = bc_type;
    free(t->bc_val[k]);
    t->bc_val[k] = NULL;
    if (b) {
        int n = t->ncell[id];
        int m = t->dim == 2 ? n : n * n;
        t->bc_val[k] = malloc((size_t)m * sizeof(double));
        memcpy(t->bc_val[k], b, (size_t)m * sizeof(double));
    }
    return 0;
}

static int n_ids(const orc_tree *t, int lvl)
{
    if (lvl < t->min_lvl || lvl > t->highest_lvl) return 0;
    return t->lvl_start[lvl - t->min_lvl + 1] - t->lvl_start[lvl - t->min_lvl];
}

static int level_id(const orc_tree *t, int lvl, int i)
{
    return t->rec_ids[t->lvl_start[lvl - t->min_lvl] + i];
}

/* a$D_has_children, m_aX_t.f90:355-358 */
static int has_children(const orc_tree *t, int id) { return t->children[id * 8] != 0; }

/* a$D_get_child_offset, m_aX_t.f90:362-374 */
static void child_offset(const orc_tree *t, int id, int off[3])
{
    off[0] = off[1] = off[2] = 0;
    if (t->lvl[id] > 1)
        for (int d = 0; d < t->dim; d++)
            off[d] = ((t->ix[id * 3 + d] - 1) & 1) * (t->ncell[id] >> 1);
}

/* --------------------------------------------------------------------- */
/* face variables                                                         */

int orc_n_face_values(const orc_tree *t) { return (int)t->n_face; }

int orc_set_fc(orc_tree *t, int ifc, const double *packed)
{
    if (ifc < 1 || ifc > ORC_MAX_FC) return -1;
    if (!t->fc[ifc - 1]) t->fc[ifc - 1] = calloc(t->n_face, sizeof(double));
    memcpy(t->fc[ifc - 1], packed, t->n_face * sizeof(double));
    return 0;
}

int orc_get_fc(const orc_tree *t, int ifc, double *packed)
{
    if (ifc < 1 || ifc > ORC_MAX_FC || !t->fc[ifc - 1]) return -1;
    memcpy(packed, t->fc[ifc - 1], t->n_face * sizeof(double));
    return 0;
}

/* Face value of direction d at 1-based cell index (x0, x1, x2) of box id:
 * box%fx / fy / fz (m_aX_core.f90:359-369), extent n+1 along d, Fortran
 * order. */
static double *face_at(const orc_tree *t, int ifc, int id, int d, int x0, int x1, int x2)
{
    int n = t->ncell[id];
    size_t len = (size_t)(n + 1) * (t->dim == 3 ? n * n : n);
    int e0 = d == 0 ? n + 1 : n, e1 = d == 1 ? n + 1 : n;
    size_t q = (size_t)(x0 - 1) + (size_t)e0 * ((x1 - 1) + (size_t)e1 * (t->dim == 3 ? x2 - 1 : 0));
    return t->fc[ifc - 1] + t->foff[id] + (size_t)d * len + q;
}

/* flux_from_children, m_aX_core.f90:838-943: the fluxes of the leaf
 * neighbour nb_id on the face it shares with parent id, set to the mean of
 * the fluxes of id's children on that face (0.5 (a + b) in 2D, 0.25
 * (((a + b) + c) + d) in 3D, first tangential index fastest).  On a
 * cylindrical tree the y faces (d = 1) take the radius-weighted mean
 * 0.5 (w1 a + w2 b) of m_aX_core.f90:877-887, with the weights of
 * a2_cyl_child_weights (m_aX_t.f90:518-528) for the neighbour's cell
 * ioff(1) + n: t = 0.25 dr / r, r = r_min(1) + (i - 0.5) dr, w = 1 -/+ t. */
static void flux_from_children(orc_tree *t, int id, int f, int ifc)
{
    int nc = t->ncell[id], nch = nc >> 1, d = f / 2, high = f % 2;
    int nb_id = t->nbr[id * 6 + f];
    int i = high ? nc + 1 : 1, i_nb = high ? 1 : nc + 1;
    int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    for (int c = 0; c < t->nch; c++) {
        if (((c >> d) & 1) != high) continue;            /* a$D_child_adj_nb */
        int c_id = t->children[id * 8 + c];
        int oa = nch * ((c >> ta) & 1), ob = t->dim == 3 ? nch * ((c >> tb) & 1) : 0;
        int nb2 = t->dim == 3 ? nch : 1;
        for (int m2 = 1; m2 <= nb2; m2++)
            for (int m = 1; m <= nch; m++) {
                int x[3], y[3];
                x[d] = i_nb; x[ta] = oa + m; x[tb] = t->dim == 3 ? ob + m2 : 1;
                double v;
                if (t->dim == 2) {
                    y[d] = i; y[ta] = 2 * m - 1; y[tb] = 1;
                    double a = *face_at(t, ifc, c_id, d, y[0], y[1], y[2]);
                    y[ta] = 2 * m;
                    double b = *face_at(t, ifc, c_id, d, y[0], y[1], y[2]);
                    if (t->cyl && d == 1) {
                        double r = t->rmin[nb_id] + (x[ta] - 0.5) * t->dr[nb_id];
                        double tw = 0.25 * t->dr[nb_id] / r;
                        double w1 = 1 - tw, w2 = 1 + tw;
                        v = 0.5 * (w1 * a + w2 * b);
                    } else {
                        v = 0.5 * (a + b);
                    }
                } else {
                    double w[4];
                    for (int k = 0; k < 4; k++) {
                        y[d] = i;
                        y[ta] = 2 * m - 1 + (k & 1);
                        y[tb] = 2 * m2 - 1 + (k >> 1);
                        w[k] = *face_at(t, ifc, c_id, d, y[0], y[1], y[2]);
                    }
                    v = 0.25 * (w[0] + w[1] + w[2] + w[3]);
                }
                *face_at(t, ifc, nb_id, d, x[0], x[1], x[2]) = v;
            }
    }
}

/* a$D_consistent_fluxes, m_aX_core.f90:811-836: every parent, lowest level
 * first, sets the fluxes of each neighbour that exists and has no
 * children. */
int orc_consistent_fluxes(orc_tree *t, int n_f, const int *f_ixs)
{
    for (int k = 0; k < n_f; k++)
        if (f_ixs[k] < 1 || f_ixs[k] > ORC_MAX_FC || !t->fc[f_ixs[k] - 1]) return -1;
    for (int lvl = t->min_lvl; lvl < t->highest_lvl; lvl++)
        for (int i = 0; i < n_ids(t, lvl); i++) {
            int id = level_id(t, lvl, i);
            if (!has_children(t, id)) continue;
            for (int f = 0; f < t->nnb; f++) {
                int nb_id = t->nbr[id * 6 + f];
                if (nb_id > 0 && !has_children(t, nb_id))
                    for (int k = 0; k < n_f; k++) flux_from_children(t, id, f, f_ixs[k]);
            }
        }
    return 0;
}

/* --------------------------------------------------------------------- */
/* ghost cells                                                            */

/* Cell (i, j, k) addressed through a permutation: a[dn] is the normal
 * coordinate of face direction dn, a[t1], a[t2] the tangential ones. */
static inline int face_ix(int n, int dn, int g, int t1v, int t2v)
{
    int a[3] = {0, 0, 0};
    a[dn] = g;
    a[dn == 0 ? 1 : 0] = t1v;
    a[dn == 2 ? 1 : 2] = t2v;
    return ix3(n, a[0], a[1], a[2]);
}

/* sides_from_nb, m_aX_gc.f90:381-415: copy the neighbour's adjacent layer. */
static void sides_from_nb(orc_tree *t, int id, int nb_id, int f, int iv)
{
    int n = t->ncell[id], dn = f / 2, high = f % 2;
    double *c = cc(t, id, iv);
    const double *s = cc(t, nb_id, iv);
    int g = high ? n + 1 : 0, src = high ? 1 : n;
    int n2 = t->dim == 3 ? n : 1;
    for (int b = 1; b <= n2; b++)
        for (int a = 1; a <= n; a++) {
            int bb = t->dim == 3 ? b : 0;
            c[face_ix(n, dn, g, a, bb)] = s[face_ix(n, dn, src, a, bb)];
        }
}

/* a$D_prolong0 restricted to one ghost layer (a$D_gc_prolong0,
 * m_aX_gc.f90:135-151; a$D_prolong0, m_aX_prolong.f90:39-108). */
static void gc_prolong0(orc_tree *t, int id, int f, int iv)
{
    int n = t->ncell[id], dn = f / 2, high = f % 2;
    int p = t->parent[id], np = t->ncell[p];
    int off[3];
    child_offset(t, id, off);
    double *c = cc(t, id, iv);
    const double *pc = cc(t, p, iv);
    int lo[3] = {1, 1, 1}, hi[3] = {n, n, n};
    if (t->dim == 2) lo[2] = hi[2] = 0;
    lo[dn] = hi[dn] = high * (n + 1);
    for (int k = lo[2]; k <= hi[2]; k++) {
        int kc = t->dim == 3 ? off[2] + ((k + 1) >> 1) : 0;
        for (int j = lo[1]; j <= hi[1]; j++) {
            int jc = off[1] + ((j + 1) >> 1);
            for (int i = lo[0]; i <= hi[0]; i++) {
                int ic = off[0] + ((i + 1) >> 1);
                c[ix3(n, i, j, k)] = pc[ix3(np, ic, jc, kc)];
            }
        }
    }
}

/* mg$D_sides_rb, m_aX_mg.f90:262-350 */
static void sides_rb(orc_tree *t, int id, int f, int iv)
{
    int n = t->ncell[id], dn = f / 2, high = f % 2;
    int ixn = high ? n : 1, dix = high ? -1 : 1;
    gc_prolong0(t, id, f, iv);
    double *c = cc(t, id, iv);
    if (t->dim == 2) {
        for (int tt = 1; tt <= n; tt++) {
            int dt = -1 + 2 * (tt & 1);
            int i, j, di, dj;
            if (dn == 0) { i = ixn; di = dix; j = tt; dj = dt; }
            else         { j = ixn; dj = dix; i = tt; di = dt; }
            int g = dn == 0 ? ix3(n, i - di, j, 0) : ix3(n, i, j - dj, 0);
            c[g] = 0.5 * c[g] + c[ix3(n, i, j, 0)] -
                   0.25 * (c[ix3(n, i + di, j, 0)] + c[ix3(n, i, j + dj, 0)]);
        }
    } else {
        for (int b = 1; b <= n; b++)
            for (int a = 1; a <= n; a++) {
                int idx[3], dd[3];
                /* tangential coordinates in increasing dimension order */
                int t1 = dn == 0 ? 1 : 0, t2 = dn == 2 ? 1 : 2;
                idx[dn] = ixn; dd[dn] = dix;
                idx[t1] = a; dd[t1] = -1 + 2 * (a & 1);
                idx[t2] = b; dd[t2] = -1 + 2 * (b & 1);
                int gi[3] = {idx[0], idx[1], idx[2]};
                gi[dn] -= dd[dn];
                int g = ix3(n, gi[0], gi[1], gi[2]);
                c[g] = 0.5 * c[g] + 0.75 * c[ix3(n, idx[0], idx[1], idx[2])] -
                       0.25 * c[ix3(n, idx[0] + dd[0], idx[1] + dd[1], idx[2] + dd[2])];
            }
    }
}

/* sides_bc (captured boundary values, e.g. a$D_bc_dirichlet_zero,
 * m_aX_gc.f90:371-379 via a$D_set_box_gc, m_aX_t.f90:436-505) followed by
 * bc_to_gc, m_aX_gc.f90:71-132. */
static void sides_bc(orc_tree *t, int id, int f, int iv)
{
    int n = t->ncell[id], dn = f / 2, high = f % 2;
    size_t k = (size_t)id * 6 + f;
    double c1, c2;
    if (t->bc_type[k] == ORC_BC_DIRICHLET) {
        c1 = -1;
        c2 = 2;
    } else {
        c1 = 1;
        c2 = t->dr[id] * (high ? 1 : -1);   /* a$D_neighb_high_pm */
    }
    const double *b = t->bc_val[k];
    double *c = cc(t, id, iv);
    int g = high ? n + 1 : 0, in = high ? n : 1;
    int n2 = t->dim == 3 ? n : 1;
    for (int bb = 1; bb <= n2; bb++)
        for (int a = 1; a <= n; a++) {
            int z = t->dim == 3 ? bb : 0;
            int gi = face_ix(n, dn, g, a, z);
            c[gi] = b ? b[(a - 1) + n * (bb - 1)] : 0.0;
            c[gi] = c2 * c[gi] + c1 * c[face_ix(n, dn, in, a, z)];
        }
}

/* a$D_gc_box, m_aX_gc.f90:50-69, with subr_rb = mg$D_sides_rb */
static void gc_box(orc_tree *t, int id, int iv)
{
    for (int f = 0; f < t->nnb; f++) {
        int nb_id = t->nbr[id * 6 + f];
        if (nb_id > 0)
            sides_from_nb(t, id, nb_id, f, iv);
        else if (nb_id == 0)
            sides_rb(t, id, f, iv);
        else
            sides_bc(t, id, f, iv);
    }
}

/* a$D_gc_interp and a$D_gc_interp_lim (lim = 1), m_aX_gc.f90:155-358:
 * ghost = interpolation between the parent's neighbour (coarse, c1..c3) and
 * the box's own first interior layer (f); the limited form caps it at
 * 2*c1.  The parent's neighbour is indexed with a$D_get_child_offset(box, nb)
 * (m_aX_t.f90:362-374), whose tangential components are the box's offset in
 * its parent; its normal index is nc (low face) or 1 (high face). */
static void gc_interp(orc_tree *t, int id, int f, int iv, int lim)
{
    const double sixth = 1 / 6.0, third = 1 / 3.0;
    int n = t->ncell[id], dn = f / 2, high = f % 2;
    int p = t->parent[id], pnb = t->nbr[p * 6 + f];
    int off[3];
    child_offset(t, id, off);
    double *c = cc(t, id, iv);
    const double *q = cc(t, pnb, iv);
    int ix = high ? n + 1 : 0, ix_f = high ? n : 1, ix_c = high ? 1 : n;
    if (t->dim == 2) {
        for (int a = 1; a <= n; a++) {
            int o = dn == 0 ? off[1] : off[0];
            int a1 = o + ((a + 1) >> 1);
            int a2 = a1 + 1 - 2 * (a & 1);
            double c1, c2, fv;
            int g;
            if (dn == 0) {              /* case (1) */
                c1 = q[ix3(n, ix_c, a1, 0)];
                c2 = q[ix3(n, ix_c, a2, 0)];
                fv = c[ix3(n, ix_f, a, 0)];
                g = ix3(n, ix, a, 0);
            } else {                    /* case (2) */
                c1 = q[ix3(n, a1, ix_c, 0)];
                c2 = q[ix3(n, a2, ix_c, 0)];
                fv = c[ix3(n, a, ix_f, 0)];
                g = ix3(n, a, ix, 0);
            }
            c[g] = 0.5 * c1 + sixth * c2 + third * fv;
            if (lim && c[g] > 2 * c1) c[g] = 2 * c1;
        }
        return;
    }
    for (int b = 1; b <= n; b++)
        for (int a = 1; a <= n; a++) {
            double c1, c2, c3, fv;
            int g;
            if (dn == 0) {              /* case (1): a = j, b = k */
                int j1 = off[1] + ((a + 1) >> 1), j2 = j1 + 1 - 2 * (a & 1);
                int k1 = off[2] + ((b + 1) >> 1), k2 = k1 + 1 - 2 * (b & 1);
                c1 = q[ix3(n, ix_c, j1, k1)];
                c2 = q[ix3(n, ix_c, j2, k1)];
                c3 = q[ix3(n, ix_c, j1, k2)];
                fv = c[ix3(n, ix_f, a, b)];
                g = ix3(n, ix, a, b);
            } else if (dn == 1) {       /* case (2): a = i, b = k */
                int i1 = off[0] + ((a + 1) >> 1), i2 = i1 + 1 - 2 * (a & 1);
                int k1 = off[2] + ((b + 1) >> 1), k2 = k1 + 1 - 2 * (b & 1);
                c1 = q[ix3(n, i1, ix_c, k1)];
                c2 = q[ix3(n, i2, ix_c, k1)];
                c3 = q[ix3(n, i1, ix_c, k2)];
                fv = c[ix3(n, a, ix_f, b)];
                g = ix3(n, a, ix, b);
            } else {                    /* case (3): a = i, b = j */
                int i1 = off[0] + ((a + 1) >> 1), i2 = i1 + 1 - 2 * (a & 1);
                int j1 = off[1] + ((b + 1) >> 1), j2 = j1 + 1 - 2 * (b & 1);
                c1 = q[ix3(n, i1, j1, ix_c)];
                c2 = q[ix3(n, i1, j2, ix_c)];
                c3 = q[ix3(n, i2, j1, ix_c)];
                fv = c[ix3(n, a, b, ix_f)];
                g = ix3(n, a, b, ix);
            }
            c[g] = third * c1 + sixth * c2 + sixth * c3 + third * fv;
            if (lim && c[g] > 2 * c1) c[g] = 2 * c1;
        }
}

/* a$D_bc_dirichlet_zero / a$D_bc_neumann_zero (m_aX_gc.f90:361-379) followed
 * by bc_to_gc: the boundary value is 0 on every face. */
static void sides_bc_zero(orc_tree *t, int id, int f, int iv, int bc_type)
{
    int n = t->ncell[id], dn = f / 2, high = f % 2;
    double c1, c2;
    if (bc_type == ORC_BC_DIRICHLET) {
        c1 = -1;
        c2 = 2;
    } else {
        c1 = 1;
        c2 = t->dr[id] * (high ? 1 : -1);
    }
    double *c = cc(t, id, iv);
    int g = high ? n + 1 : 0, in = high ? n : 1;
    int n2 = t->dim == 3 ? n : 1;
    for (int bb = 1; bb <= n2; bb++)
        for (int a = 1; a <= n; a++) {
            int z = t->dim == 3 ? bb : 0;
            int gi = face_ix(n, dn, g, a, z);
            c[gi] = 0.0;
            c[gi] = c2 * c[gi] + c1 * c[face_ix(n, dn, in, a, z)];
        }
}

/* a$D_gc_tree, m_aX_gc.f90:28-69: every box of every level, lowest level
 * first, with subr_rb = mg$D_sides_rb / a$D_gc_prolong0 / a$D_gc_interp /
 * a$D_gc_interp_lim (rb = ORC_RB_*) and subr_bc = the boundary table
 * (bc = 0) or a$D_bc_dirichlet_zero / a$D_bc_neumann_zero (bc =
 * ORC_BC_DIRICHLET / ORC_BC_NEUMANN). */
int orc_gc_tree(orc_tree *t, int iv, int rb, int bc)
{
    if (rb < ORC_RB_MG || rb > ORC_RB_INTERP_LIM) return 1;
    if (bc != 0 && bc != ORC_BC_DIRICHLET && bc != ORC_BC_NEUMANN) return 1;
    for (int lvl = t->min_lvl; lvl <= t->highest_lvl; lvl++)
        for (int i = 0; i < n_ids(t, lvl); i++) {
            int id = level_id(t, lvl, i);
            for (int f = 0; f < t->nnb; f++) {
                int nb_id = t->nbr[id * 6 + f];
                if (nb_id > 0)
                    sides_from_nb(t, id, nb_id, f, iv);
                else if (nb_id == 0) {
                    if (rb == ORC_RB_MG) sides_rb(t, id, f, iv);
                    else if (rb == ORC_RB_PROLONG0) gc_prolong0(t, id, f, iv);
                    else gc_interp(t, id, f, iv, rb == ORC_RB_INTERP_LIM);
                } else if (bc == 0)
                    sides_bc(t, id, f, iv);
                else
                    sides_bc_zero(t, id, f, iv, bc);
            }
        }
    return 0;
}

/* fill_gc_phi, m_aX_mg.f90:352-365 (over all boxes of a level) */
int orc_fill_gc_level(orc_tree *t, int iv, int lvl)
{
    for (int i = 0; i < n_ids(t, lvl); i++) gc_box(t, level_id(t, lvl, i), iv);
    return 0;
}

/* --------------------------------------------------------------------- */
/* transfer operators                                                     */

/* a$D_restrict_box, m_aX_restrict.f90:67-132 (Cartesian branch) */
static void restrict_box(orc_tree *t, int c_id, int p_id, int iv, int iv_to)
{
    int n = t->ncell[c_id], np = t->ncell[p_id], hnc = n >> 1;
    int off[3];
    child_offset(t, c_id, off);
    const double *c = cc(t, c_id, iv);
    double *p = cc(t, p_id, iv_to);
    if (t->dim == 2 && t->cyl) {
        /* cylindrical branch, m_aX_restrict.f90:90-103; the weights are those
         * of the parent's cell i (not i_c), a2_cyl_child_weights
         * (m_aX_t.f90:519-528) */
        for (int j = 1; j <= hnc; j++)
            for (int i = 1; i <= hnc; i++) {
                int f = 2 * i - 1, g = 2 * j - 1;
                double r = t->rmin[p_id] + (i - 0.5) * t->dr[p_id];
                double tmp = 0.25 * t->dr[p_id] / r, w1 = 1 - tmp, w2 = 1 + tmp;
                p[ix3(np, off[0] + i, off[1] + j, 0)] =
                    0.25 * (w1 * (c[ix3(n, f, g, 0)] + c[ix3(n, f, g + 1, 0)]) +
                            w2 * (c[ix3(n, f + 1, g, 0)] + c[ix3(n, f + 1, g + 1, 0)]));
            }
    } else if (t->dim == 2) {
        for (int j = 1; j <= hnc; j++)
            for (int i = 1; i <= hnc; i++) {
                int f = 2 * i - 1, g = 2 * j - 1;
                double s = c[ix3(n, f, g, 0)];
                s += c[ix3(n, f + 1, g, 0)];
                s += c[ix3(n, f, g + 1, 0)];
                s += c[ix3(n, f + 1, g + 1, 0)];
                p[ix3(np, off[0] + i, off[1] + j, 0)] = 0.25 * s;
            }
    } else {
        for (int k = 1; k <= hnc; k++)
            for (int j = 1; j <= hnc; j++)
                for (int i = 1; i <= hnc; i++) {
                    int f = 2 * i - 1, g = 2 * j - 1, h = 2 * k - 1;
                    double s = c[ix3(n, f, g, h)];
                    s += c[ix3(n, f + 1, g, h)];
                    s += c[ix3(n, f, g + 1, h)];
                    s += c[ix3(n, f + 1, g + 1, h)];
                    s += c[ix3(n, f, g, h + 1)];
                    s += c[ix3(n, f + 1, g, h + 1)];
                    s += c[ix3(n, f, g + 1, h + 1)];
                    s += c[ix3(n, f + 1, g + 1, h + 1)];
                    p[ix3(np, off[0] + i, off[1] + j, off[2] + k)] = 0.125 * s;
                }
    }
}

/* a$D_prolong2, m_aX_prolong.f90:243-348 */
static void prolong2(orc_tree *t, int p_id, int c_id, int iv, int ivc, int add)
{
    int n = t->ncell[c_id], np = t->ncell[p_id], hnc = n >> 1;
    int off[3];
    child_offset(t, c_id, off);
    const double *p = cc(t, p_id, iv);
    double *c = cc(t, c_id, ivc);
    if (!add) {
        int kk = t->dim == 3 ? n : 0;
        for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
            for (int j = 1; j <= n; j++)
                for (int i = 1; i <= n; i++) c[ix3(n, i, j, k)] = 0;
    }
    if (t->dim == 2) {
        for (int j = 1; j <= hnc; j++) {
            int jc = j + off[1], jf = 2 * j - 1;
            for (int i = 1; i <= hnc; i++) {
                int ic = i + off[0], i_f = 2 * i - 1;
                double f0 = p[ix3(np, ic, jc, 0)];
                double fx = 0.125 * (p[ix3(np, ic + 1, jc, 0)] - p[ix3(np, ic - 1, jc, 0)]);
                double fy = 0.125 * (p[ix3(np, ic, jc + 1, 0)] - p[ix3(np, ic, jc - 1, 0)]);
                double fxx = 0.03125 * (p[ix3(np, ic - 1, jc, 0)] - 2 * f0 + p[ix3(np, ic + 1, jc, 0)]);
                double fyy = 0.03125 * (p[ix3(np, ic, jc - 1, 0)] - 2 * f0 + p[ix3(np, ic, jc + 1, 0)]);
                double f2 = fxx + fyy;
                double *q;
                q = &c[ix3(n, i_f, jf, 0)];         *q = f0 - fx - fy + f2 + *q;
                q = &c[ix3(n, i_f + 1, jf, 0)];     *q = f0 + fx - fy + f2 + *q;
                q = &c[ix3(n, i_f, jf + 1, 0)];     *q = f0 - fx + fy + f2 + *q;
                q = &c[ix3(n, i_f + 1, jf + 1, 0)]; *q = f0 + fx + fy + f2 + *q;
            }
        }
    } else {
        for (int k = 1; k <= hnc; k++) {
            int kc = k + off[2], kf = 2 * k - 1;
            for (int j = 1; j <= hnc; j++) {
                int jc = j + off[1], jf = 2 * j - 1;
                for (int i = 1; i <= hnc; i++) {
                    int ic = i + off[0], i_f = 2 * i - 1;
                    double f0 = p[ix3(np, ic, jc, kc)];
                    double fx = 0.125 * (p[ix3(np, ic + 1, jc, kc)] - p[ix3(np, ic - 1, jc, kc)]);
                    double fy = 0.125 * (p[ix3(np, ic, jc + 1, kc)] - p[ix3(np, ic, jc - 1, kc)]);
                    double fz = 0.125 * (p[ix3(np, ic, jc, kc + 1)] - p[ix3(np, ic, jc, kc - 1)]);
                    double fxx = 0.03125 * (p[ix3(np, ic - 1, jc, kc)] - 2 * f0 + p[ix3(np, ic + 1, jc, kc)]);
                    double fyy = 0.03125 * (p[ix3(np, ic, jc - 1, kc)] - 2 * f0 + p[ix3(np, ic, jc + 1, kc)]);
                    double fzz = 0.03125 * (p[ix3(np, ic, jc, kc - 1)] - 2 * f0 + p[ix3(np, ic, jc, kc + 1)]);
                    double f2 = fxx + fyy + fzz;
                    for (int s = 0; s < 8; s++) {
                        int sx = s & 1, sy = (s >> 1) & 1, sz = (s >> 2) & 1;
                        double v = sx ? f0 + fx : f0 - fx;
                        v = sy ? v + fy : v - fy;
                        v = sz ? v + fz : v - fz;
                        double *q = &c[ix3(n, i_f + sx, jf + sy, kf + sz)];
                        *q = v + f2 + *q;
                    }
                }
            }
        }
    }
}

/* --------------------------------------------------------------------- */
/* operators                                                              */

/* minval and maxval of variable iv over the whole box array, ghosts included */
static void box_min_max(const orc_tree *t, int id, int iv, double *lo, double *hi)
{
    const double *e = cc(t, id, iv);
    int len = box_len(t, t->ncell[id]);
    double a = e[0], b = e[0];
    for (int q = 1; q < len; q++) {
        if (e[q] < a) a = e[q];
        if (e[q] > b) b = e[q];
    }
    *lo = a;
    *hi = b;
}

/* mg$D_set_box_tag, m_aX_mg.f90:667-706 */
static void set_box_tag(orc_tree *t, int id, const orc_mg_opts *o)
{
    int is_lsf = 0, is_eps = 0, is_deps = 0;
    double a, b;
    if (o->i_lsf != -1) {   /* :680-688 */
        box_min_max(t, id, o->i_lsf, &a, &b);
        is_lsf = a * b < 0;
    }
    if (o->i_eps != -1) {   /* :690-698 */
        box_min_max(t, id, o->i_eps, &a, &b);
        is_deps = b > a;
        if (!is_deps) is_eps = (a < 1 || a > 1);
    }
    if (is_lsf + is_eps + is_deps > 1) t->tag_err = 1;   /* :700-701 stops */
    t->tag[id] = TAG_NORMAL;
    if (is_lsf) t->tag[id] = TAG_LSF;
    if (is_eps) t->tag[id] = TAG_CEPS;
    if (is_deps) t->tag[id] = TAG_VEPS;
}

/* the box's tag under the auto operators, set lazily (m_aX_mg.f90:567,602,655) */
static int auto_tag(orc_tree *t, int id, const orc_mg_opts *o)
{
    if (t->tag[id] == TAG_UNSET) set_box_tag(t, id, o);
    return t->tag[id];
}

static int use_lpld(orc_tree *t, int id, const orc_mg_opts *o)
{
    if (o->op != 1) return 0;
    int tg = auto_tag(t, id, o);
    return tg == TAG_CEPS || tg == TAG_VEPS;
}

/* the lpllsf operators for this box: every box with the lpllsf set, boxes
 * tagged mg_lsf_box with the auto set */
static int use_lsf(orc_tree *t, int id, const orc_mg_opts *o)
{
    if (o->op == 2) return 1;
    if (o->op != 1) return 0;
    return auto_tag(t, id, o) == TAG_LSF;
}

/* mg$D_box_gsrb_lpl, m_aX_mg.f90:718-760 */
static void gsrb_lpl(orc_tree *t, int id, int rb, const orc_mg_opts *o)
{
    int n = t->ncell[id];
    double dx2 = t->dr[id] * t->dr[id];
    double *p = cc(t, id, o->i_phi);
    const double *r = cc(t, id, o->i_rhs);
    const double sixth = 1 / 6.0;
    if (t->dim == 2) {
        for (int j = 1; j <= n; j++) {
            int i0 = 2 - ((rb ^ j) & 1);
            for (int i = i0; i <= n; i += 2) {
                int q = ix3(n, i, j, 0);
                p[q] = 0.25 * (p[q + 1] + p[q - 1] + p[q + (n + 2)] + p[q - (n + 2)] -
                               dx2 * r[q]);
            }
        }
    } else {
        int s = n + 2, s2 = s * s;
        for (int k = 1; k <= n; k++)
            for (int j = 1; j <= n; j++) {
                int i0 = 2 - ((rb ^ (k + j)) & 1);
                for (int i = i0; i <= n; i += 2) {
                    int q = ix3(n, i, j, k);
                    p[q] = sixth * (p[q + 1] + p[q - 1] + p[q + s] + p[q - s] + p[q + s2] +
                                    p[q - s2] - dx2 * r[q]);
                }
            }
    }
}

/* mg$D_box_gsrb_lpld, m_aX_mg.f90:810-863 */
static void gsrb_lpld(orc_tree *t, int id, int rb, const orc_mg_opts *o)
{
    int n = t->ncell[id];
    double dx2 = t->dr[id] * t->dr[id];
    double *p = cc(t, id, o->i_phi);
    const double *r = cc(t, id, o->i_rhs);
    const double *e = cc(t, id, o->i_eps);
    int s = n + 2, s2 = s * s;
    int nn = 2 * t->dim;
    int dq[6] = {-1, 1, -s, s, -s2, s2};
    int kk = t->dim == 3 ? n : 0;
    for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
        for (int j = 1; j <= n; j++) {
            int i0 = 2 - ((rb ^ (t->dim == 3 ? k + j : j)) & 1);
            for (int i = i0; i <= n; i += 2) {
                int q = ix3(n, i, j, k);
                double a0 = e[q], c[6] = {0, 0, 0, 0, 0, 0}, su, sc;
                for (int m = 0; m < nn; m++) c[m] = 2 * a0 * e[q + dq[m]] / (a0 + e[q + dq[m]]);
                su = c[0] * p[q + dq[0]];
                for (int m = 1; m < nn; m++) su += c[m] * p[q + dq[m]];
                sc = c[0];
                for (int m = 1; m < nn; m++) sc += c[m];
                p[q] = (su - dx2 * r[q]) / sc;
            }
        }
}

/* mg$D_box_lpl, m_aX_mg.f90:763-797 */
static void box_lpl(orc_tree *t, int id, int i_out, const orc_mg_opts *o)
{
    int n = t->ncell[id];
    double inv_dr_sq = 1 / (t->dr[id] * t->dr[id]);
    const double *p = cc(t, id, o->i_phi);
    double *out = cc(t, id, i_out);
    int s = n + 2, s2 = s * s;
    if (t->dim == 2) {
        for (int j = 1; j <= n; j++)
            for (int i = 1; i <= n; i++) {
                int q = ix3(n, i, j, 0);
                out[q] = inv_dr_sq * (p[q - 1] + p[q + 1] + p[q - s] + p[q + s] - 4 * p[q]);
            }
    } else {
        for (int k = 1; k <= n; k++)
            for (int j = 1; j <= n; j++)
                for (int i = 1; i <= n; i++) {
                    int q = ix3(n, i, j, k);
                    out[q] = inv_dr_sq * (p[q - 1] + p[q + 1] + p[q - s] + p[q + s] +
                                          p[q - s2] + p[q + s2] - 6 * p[q]);
                }
    }
}

/* mg$D_box_lpld, m_aX_mg.f90:866-916 */
static void box_lpld(orc_tree *t, int id, int i_out, const orc_mg_opts *o)
{
    int n = t->ncell[id];
    double inv_dr_sq = 1 / (t->dr[id] * t->dr[id]);
    const double *p = cc(t, id, o->i_phi);
    const double *e = cc(t, id, o->i_eps);
    double *out = cc(t, id, i_out);
    int s = n + 2, s2 = s * s;
    int nn = 2 * t->dim;
    int dq[6] = {-1, 1, -s, s, -s2, s2};
    int kk = t->dim == 3 ? n : 0;
    for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
        for (int j = 1; j <= n; j++)
            for (int i = 1; i <= n; i++) {
                int q = ix3(n, i, j, k);
                double u0 = p[q], a0 = e[q], sum;
                double a = e[q + dq[0]];
                sum = a0 * a / (a0 + a) * (p[q + dq[0]] - u0);
                for (int m = 1; m < nn; m++) {
                    a = e[q + dq[m]];
                    sum += a0 * a / (a0 + a) * (p[q + dq[m]] - u0);
                }
                out[q] = inv_dr_sq * 2 * sum;
            }
}

/* mg$D_box_corr_lpld, m_aX_mg.f90:918-985 */
static void corr_lpld(orc_tree *t, int p_id, int c_id, const orc_mg_opts *o)
{
    int n = t->ncell[c_id], np = t->ncell[p_id];
    int off[3];
    child_offset(t, c_id, off);
    const double *pc = cc(t, p_id, o->i_tmp);
    const double *pe = cc(t, p_id, o->i_eps);
    double *c = cc(t, c_id, o->i_phi);
    const double third = 1 / 3.0;
    int kk = t->dim == 3 ? n : 0;
    for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++) {
        int kc1 = 0, kc2 = 0;
        if (t->dim == 3) {
            kc1 = off[2] + ((k + 1) >> 1);
            kc2 = kc1 + 1 - 2 * (k & 1);
        }
        for (int j = 1; j <= n; j++) {
            int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
            for (int i = 1; i <= n; i++) {
                int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
                int q0 = ix3(np, ic1, jc1, kc1);
                int qn[3] = {ix3(np, ic2, jc1, kc1), ix3(np, ic1, jc2, kc1), ix3(np, ic1, jc1, kc2)};
                double u0 = pc[q0], a0 = pe[q0];
                double *q = &c[ix3(n, i, j, k)];
                if (t->dim == 2) {
                    double s = (a0 * u0 + pe[qn[0]] * pc[qn[0]]) / (a0 + pe[qn[0]]);
                    s += (a0 * u0 + pe[qn[1]] * pc[qn[1]]) / (a0 + pe[qn[1]]);
                    *q = *q + 0.5 * s;
                } else {
                    double s = (a0 * u0 + pe[qn[0]] * (1.5 * pc[qn[0]] - 0.5 * u0)) / (a0 + pe[qn[0]]);
                    for (int m = 1; m < 3; m++)
                        s += (a0 * u0 + pe[qn[m]] * (1.5 * pc[qn[m]] - 0.5 * u0)) / (a0 + pe[qn[m]]);
                    *q = *q + third * s;
                }
            }
        }
    }
}

/* --------------------------------------------------------------------- */
/* cylindrical coordinates (2D, x the radius)                             */

/* box%r_min(1) of every box as the reference computes it: r_base +
 * (ix-1) dr_base n_cell on levels <= 1 (m_aX_core.f90:195-196), the parent's
 * plus 0.5 dr_p dix n_cell for a child (:774-775). */
int orc_set_coord(orc_tree *t, int cyl, double r_base)
{
    if (cyl && t->dim != 2) return -1;
    t->cyl = cyl;
    free(t->rmin);
    t->rmin = NULL;
    if (!cyl) return 0;
    t->rmin = calloc((size_t)t->max_id + 1, sizeof(double));
    for (int r = 0; r < t->n_boxes; r++) {   /* level order: parents first */
        int id = t->rec_ids[r], ix = t->ix[id * 3];
        if (t->lvl[id] <= 1)
            t->rmin[id] = r_base + (ix - 1) * t->dr_base * t->nc;
        else {
            int p = t->parent[id];
            t->rmin[id] = t->rmin[p] + 0.5 * t->dr[p] * ((ix - 1) & 1) * t->ncell[p];
        }
    }
    return 0;
}

/* rfac = [i+ioff-1, i+ioff] / (i+ioff-0.5), ioff = (ix(1)-1) n_cell
 * (m_aX_mg.f90:1338,1345) */
static void cyl_rfac(const orc_tree *t, int id, int i, double rf[2])
{
    int ir = i + (t->ix[id * 3] - 1) * t->ncell[id];
    double d = ir - 0.5;
    rf[0] = (ir - 1) / d;
    rf[1] = ir / d;
}

/* mg2_box_gsrb_clpl, m_aX_mg.f90:1325-1351 */
static void gsrb_clpl(orc_tree *t, int id, int rb, const orc_mg_opts *o)
{
    int n = t->ncell[id], s = n + 2;
    double dx2 = t->dr[id] * t->dr[id];
    double *p = cc(t, id, o->i_phi);
    const double *r = cc(t, id, o->i_rhs);
    for (int j = 1; j <= n; j++) {
        int i0 = 2 - ((rb ^ j) & 1);
        for (int i = i0; i <= n; i += 2) {
            int q = ix3(n, i, j, 0);
            double rf[2];
            cyl_rfac(t, id, i, rf);
            p[q] = 0.25 * (rf[0] * p[q - 1] + rf[1] * p[q + 1] + p[q + s] + p[q - s] - dx2 * r[q]);
        }
    }
}

/* mg2_box_gsrb_clpld, m_aX_mg.f90:1411-1443 */
static void gsrb_clpld(orc_tree *t, int id, int rb, const orc_mg_opts *o)
{
    int n = t->ncell[id], s = n + 2;
    double dx2 = t->dr[id] * t->dr[id];
    double *p = cc(t, id, o->i_phi);
    const double *r = cc(t, id, o->i_rhs);
    const double *e = cc(t, id, o->i_eps);
    int dq[4] = {-1, 1, -s, s};
    for (int j = 1; j <= n; j++) {
        int i0 = 2 - ((rb ^ j) & 1);
        for (int i = i0; i <= n; i += 2) {
            int q = ix3(n, i, j, 0);
            double rf[4] = {0, 0, 1, 1}, c[4], su, sc;
            cyl_rfac(t, id, i, rf);
            double a0 = e[q];
            for (int m = 0; m < 4; m++) c[m] = 2 * a0 * e[q + dq[m]] / (a0 + e[q + dq[m]]);
            su = c[0] * rf[0] * p[q + dq[0]];
            for (int m = 1; m < 4; m++) su += c[m] * rf[m] * p[q + dq[m]];
            sc = c[0] * rf[0];
            for (int m = 1; m < 4; m++) sc += c[m] * rf[m];
            p[q] = (su - dx2 * r[q]) / sc;
        }
    }
}

/* mg2_box_clpl, m_aX_mg.f90:1353-1376 */
static void box_clpl(orc_tree *t, int id, int i_out, const orc_mg_opts *o)
{
    int n = t->ncell[id], s = n + 2;
    double inv_dr_sq = 1 / (t->dr[id] * t->dr[id]);
    const double *p = cc(t, id, o->i_phi);
    double *out = cc(t, id, i_out);
    for (int j = 1; j <= n; j++)
        for (int i = 1; i <= n; i++) {
            int q = ix3(n, i, j, 0);
            double rf[2];
            cyl_rfac(t, id, i, rf);
            out[q] = (rf[0] * p[q - 1] + rf[1] * p[q + 1] + p[q - s] + p[q + s] - 4 * p[q]) * inv_dr_sq;
        }
}

/* mg2_box_clpld, m_aX_mg.f90:1378-1408 */
static void box_clpld(orc_tree *t, int id, int i_out, const orc_mg_opts *o)
{
    int n = t->ncell[id], s = n + 2;
    double inv_dr_sq = 1 / (t->dr[id] * t->dr[id]);
    const double *p = cc(t, id, o->i_phi);
    const double *e = cc(t, id, o->i_eps);
    double *out = cc(t, id, i_out);
    int dq[4] = {-1, 1, -s, s};
    for (int j = 1; j <= n; j++)
        for (int i = 1; i <= n; i++) {
            int q = ix3(n, i, j, 0);
            double rf[4] = {0, 0, 1, 1}, u0 = p[q], a0 = e[q], sum = 0;
            cyl_rfac(t, id, i, rf);
            for (int m = 0; m < 4; m++) {
                double a = e[q + dq[m]];
                double term = rf[m] * a0 * a / (a0 + a) * (p[q + dq[m]] - u0);
                sum = m == 0 ? term : sum + term;
            }
            out[q] = inv_dr_sq * 2 * sum;
        }
}

/* --------------------------------------------------------------------- */
/* internal boundaries given by a level set (m_aX_mg.f90:987-1321)         */

/* lsf_dist_val, m_aX_mg.f90:990-1017: distance (0..1] to the neighbour b and
 * the value there, the interface's boundary value where the level set
 * changes sign between a and b */
static void lsf_dist_val(double la, double ba, double lb, double vb, double bb, double *dist,
                         double *val)
{
    if (la * lb < 0) {
        double d = la / (la - lb);
        *dist = d;
        *val = ba * (1 - d) + bb * d;
    } else {
        *dist = 1;
        *val = vb;
    }
}

/* distances and values of the 2*dim face neighbours of cell q, in the order
 * i-1, i+1, j-1, j+1 (, k-1, k+1) */
static void lsf_neighbours(const orc_tree *t, int id, const orc_mg_opts *o, int q, double dd[7],
                           double val[7])
{
    int n = t->ncell[id], s = n + 2, s2 = s * s;
    const double *p = cc(t, id, o->i_phi), *l = cc(t, id, o->i_lsf), *b = cc(t, id, o->i_bval);
    int dq[6] = {-1, 1, -s, s, -s2, s2};
    for (int m = 0; m < 2 * t->dim; m++) {
        int qb = q + dq[m];
        lsf_dist_val(l[q], b[q], l[qb], p[qb], b[qb], &dd[m + 1], &val[m + 1]);
    }
}

/* mg$D_box_gsrb_lpllsf, m_aX_mg.f90:1105-1179 */
static void gsrb_lpllsf(orc_tree *t, int id, int rb, const orc_mg_opts *o)
{
    int n = t->ncell[id];
    double dx2 = t->dr[id] * t->dr[id];
    double *p = cc(t, id, o->i_phi);
    const double *r = cc(t, id, o->i_rhs);
    double d[7], v[7];
    int kk = t->dim == 3 ? n : 0;
    for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
        for (int j = 1; j <= n; j++) {
            int i0 = 2 - ((rb ^ (t->dim == 3 ? k + j : j)) & 1);
            for (int i = i0; i <= n; i += 2) {
                int q = ix3(n, i, j, k);
                lsf_neighbours(t, id, o, q, d, v);
                if (t->dim == 2) {
                    double prod = d[1] * d[2] * d[3] * d[4];
                    p[q] = 1 / (d[1] * d[2] + d[3] * d[4]) *
                           ((d[2] * v[1] + d[1] * v[2]) * d[3] * d[4] / (d[1] + d[2]) +
                            (d[4] * v[3] + d[3] * v[4]) * d[1] * d[2] / (d[3] + d[4]) -
                            0.5 * prod * dx2 * r[q]);
                } else {
                    p[q] = 1 / (1 / (d[1] * d[2]) + 1 / (d[3] * d[4]) + 1 / (d[5] * d[6])) *
                           ((d[2] * v[1] + d[1] * v[2]) / ((d[1] + d[2]) * d[1] * d[2]) +
                            (d[4] * v[3] + d[3] * v[4]) / ((d[3] + d[4]) * d[3] * d[4]) +
                            (d[6] * v[5] + d[5] * v[6]) / ((d[5] + d[6]) * d[5] * d[6]) -
                            0.5 * dx2 * r[q]);
                }
            }
        }
}

/* mg$D_box_lpllsf, m_aX_mg.f90:1182-1251 */
static void box_lpllsf(orc_tree *t, int id, int i_out, const orc_mg_opts *o)
{
    int n = t->ncell[id];
    double inv_dr_sq = 1 / (t->dr[id] * t->dr[id]);
    const double *p = cc(t, id, o->i_phi);
    double *out = cc(t, id, i_out);
    double d[7], v[7];
    int kk = t->dim == 3 ? n : 0;
    for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
        for (int j = 1; j <= n; j++)
            for (int i = 1; i <= n; i++) {
                int q = ix3(n, i, j, k);
                lsf_neighbours(t, id, o, q, d, v);
                double f0 = p[q];
                double sum = (d[2] * v[1] + d[1] * v[2] - (d[1] + d[2]) * f0) /
                                 ((d[1] + d[2]) * d[1] * d[2]) +
                             (d[4] * v[3] + d[3] * v[4] - (d[3] + d[4]) * f0) /
                                 ((d[3] + d[4]) * d[3] * d[4]);
                if (t->dim == 3)
                    sum = sum + (d[6] * v[5] + d[5] * v[6] - (d[5] + d[6]) * f0) /
                                    ((d[5] + d[6]) * d[5] * d[6]);
                out[q] = 2 * inv_dr_sq * sum;
            }
}

/* mg$D_box_corr_lpllsf, m_aX_mg.f90:1019-1102: the correction (the parent's
 * tmp) interpolated from the closest coarse cell and the one-but-closest in
 * each direction, with boundary value 0 where the level set changes sign */
static void corr_lpllsf(orc_tree *t, int p_id, int c_id, const orc_mg_opts *o)
{
    int n = t->ncell[c_id], np = t->ncell[p_id];
    int off[3];
    child_offset(t, c_id, off);
    double *ph = cc(t, c_id, o->i_phi);
    const double *lc = cc(t, c_id, o->i_lsf);
    const double *lp = cc(t, p_id, o->i_lsf), *cp = cc(t, p_id, o->i_tmp);
    int kk = t->dim == 3 ? n : 0;
    for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
        for (int j = 1; j <= n; j++)
            for (int i = 1; i <= n; i++) {
                int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
                int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
                int kc1 = 0, kc2 = 0;
                if (t->dim == 3) {
                    kc1 = off[2] + ((k + 1) >> 1);
                    kc2 = kc1 + 1 - 2 * (k & 1);
                }
                int q = ix3(n, i, j, k);
                int qb[4] = {ix3(np, ic1, jc1, kc1), ix3(np, ic2, jc1, kc1), ix3(np, ic1, jc2, kc1),
                             ix3(np, ic1, jc1, kc2)};
                double d[5] = {0}, v[5] = {0}, c[5] = {0};
                for (int m = 0; m <= t->dim; m++)
                    lsf_dist_val(lc[q], 0.0, lp[qb[m]], cp[qb[m]], 0.0, &d[m + 1], &v[m + 1]);
                double sv, sc;
                if (t->dim == 2) {
                    c[1] = 2 * d[2] * d[3];
                    c[2] = d[1] * d[3];
                    c[3] = d[1] * d[2];
                } else {
                    c[1] = d[2] * d[3] * d[4];
                    c[2] = d[1] * d[3] * d[4];
                    c[3] = d[1] * d[2] * d[4];
                    c[4] = d[1] * d[2] * d[3];
                }
                /* sum(c * val) / sum(c): Fortran SUM starts from zero */
                sv = 0.0;
                sc = 0.0;
                for (int m = 1; m <= t->dim + 1; m++) {
                    sv += c[m] * v[m];
                    sc += c[m];
                }
                ph[q] = ph[q] + sv / sc;
            }
}

/* mg$D_box_rstr_lpllsf, m_aX_mg.f90:1254-1320: children on the parent cell's
 * side of the interface only, where some but not all are.  The factor
 * 1 / n_ch is integer division in the reference, so a mean over 2..2^dim-1
 * children comes out as 0 * sum. */
static void rstr_lpllsf(orc_tree *t, int c_id, int p_id, int iv, const orc_mg_opts *o)
{
    int n = t->ncell[c_id], np = t->ncell[p_id], hnc = n >> 1;
    int off[3];
    child_offset(t, c_id, off);
    const double *c = cc(t, c_id, iv), *lc = cc(t, c_id, o->i_lsf);
    double *p = cc(t, p_id, iv);
    const double *lp = cc(t, p_id, o->i_lsf);
    int kk = t->dim == 3 ? hnc : 0;
    for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
        for (int j = 1; j <= hnc; j++)
            for (int i = 1; i <= hnc; i++) {
                int pq = ix3(np, off[0] + i, off[1] + j, t->dim == 3 ? off[2] + k : 0);
                int f = 2 * i - 1, g = 2 * j - 1, h = t->dim == 3 ? 2 * k - 1 : 0;
                int qs[8] = {0}, mask[8] = {0}, n_ch = 0;
                for (int m = 0; m < t->nch; m++) {
                    qs[m] = ix3(n, f + (m & 1), g + ((m >> 1) & 1), h + ((m >> 2) & 1));
                    mask[m] = lp[pq] * lc[qs[m]] > 0;
                    n_ch += mask[m];
                }
                if (n_ch < t->nch && n_ch > 0) {
                    double sm = 0.0;
                    for (int m = 0; m < t->nch; m++)
                        if (mask[m]) sm += c[qs[m]];
                    p[pq] = (double)(1 / n_ch) * sm;
                } else {
                    double sm = c[qs[0]];
                    for (int m = 1; m < t->nch; m++) sm += c[qs[m]];
                    p[pq] = (t->dim == 2 ? 0.25 : 0.125) * sm;
                }
            }
}

/* mg%box_rstr: mg$D_box_rstr_lpl, mg$D_auto_rstr (m_aX_mg.f90:629-644; the
 * child's tag must be set) or mg$D_box_rstr_lpllsf */
static void box_rstr(orc_tree *t, int c_id, int p_id, int iv, const orc_mg_opts *o)
{
    int lsf = o->rstr == 2;
    if (o->rstr == 1) {
        if (t->tag[c_id] == TAG_UNSET) t->tag_err = 1;   /* :639 stops */
        lsf = t->tag[c_id] == TAG_LSF;
    }
    if (lsf)
        rstr_lpllsf(t, c_id, p_id, iv, o);
    else
        restrict_box(t, c_id, p_id, iv, iv);
}

/* every box's tag, as a program calling mg$D_set_box_tag on each box does */
int orc_set_tags(orc_tree *t, const orc_mg_opts *o)
{
    for (int id = 1; id <= t->max_id; id++)
        if (t->ncell[id] > 0) set_box_tag(t, id, o);
    return t->tag_err ? -1 : 0;
}

/* mg%box_gsrb: mg$D_box_gsrb_lpl or mg$D_auto_gsrb (m_aX_mg.f90:562-590) */
static void box_gsrb(orc_tree *t, int id, int rb, const orc_mg_opts *o)
{
    if (use_lsf(t, id, o)) {
        gsrb_lpllsf(t, id, rb, o);
        return;
    }
    if (t->cyl && o->op == 1) {
        if (use_lpld(t, id, o)) gsrb_clpld(t, id, rb, o);
        else gsrb_clpl(t, id, rb, o);
        return;
    }
    if (use_lpld(t, id, o))
        gsrb_lpld(t, id, rb, o);
    else
        gsrb_lpl(t, id, rb, o);
}

/* mg%box_op: mg$D_box_lpl or mg$D_auto_op (m_aX_mg.f90:593-627) */
static void box_op(orc_tree *t, int id, int i_out, const orc_mg_opts *o)
{
    if (use_lsf(t, id, o)) {
        box_lpllsf(t, id, i_out, o);
        return;
    }
    if (t->cyl && o->op == 1) {
        if (use_lpld(t, id, o)) box_clpld(t, id, i_out, o);
        else box_clpl(t, id, i_out, o);
        return;
    }
    if (use_lpld(t, id, o))
        box_lpld(t, id, i_out, o);
    else
        box_lpl(t, id, i_out, o);
}

/* mg%box_corr: mg$D_box_corr_lpl (m_aX_mg.f90:708-715) or mg$D_auto_corr
 * (m_aX_mg.f90:646-665); the tag is that of the child. */
static void box_corr(orc_tree *t, int p_id, int c_id, const orc_mg_opts *o)
{
    if (use_lsf(t, c_id, o))
        corr_lpllsf(t, p_id, c_id, o);
    else if (use_lpld(t, c_id, o))
        corr_lpld(t, p_id, c_id, o);
    else
        prolong2(t, p_id, c_id, o->i_tmp, o->i_phi, 1);
}

/* residual_box, m_aX_mg.f90:545-559 */
static void residual_box(orc_tree *t, int id, const orc_mg_opts *o)
{
    box_op(t, id, o->i_tmp, o);
    int n = t->ncell[id];
    double *tm = cc(t, id, o->i_tmp);
    const double *r = cc(t, id, o->i_rhs);
    int kk = t->dim == 3 ? n : 0;
    for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
        for (int j = 1; j <= n; j++)
            for (int i = 1; i <= n; i++) {
                int q = ix3(n, i, j, k);
                tm[q] = r[q] - tm[q];
            }
}

/* --------------------------------------------------------------------- */
/* multigrid stages                                                       */

/* gsrb_boxes, m_aX_mg.f90:397-420 */
int orc_gsrb_level(orc_tree *t, const orc_mg_opts *o, int lvl, int n_cycle)
{
    int nb = n_ids(t, lvl);
    for (int rb = 1; rb <= 2 * n_cycle; rb++) {
        for (int i = 0; i < nb; i++) box_gsrb(t, level_id(t, lvl, i), rb, o);
        for (int i = 0; i < nb; i++) gc_box(t, level_id(t, lvl, i), o->i_phi);
    }
    return 0;
}

int orc_residual_level(orc_tree *t, const orc_mg_opts *o, int lvl)
{
    for (int i = 0; i < n_ids(t, lvl); i++) residual_box(t, level_id(t, lvl, i), o);
    return 0;
}

int orc_restrict_level(orc_tree *t, int iv, int lvl)
{
    for (int i = 0; i < n_ids(t, lvl); i++) {
        int id = level_id(t, lvl, i);
        restrict_box(t, id, t->parent[id], iv, iv);
    }
    return 0;
}

int orc_prolong2_level(orc_tree *t, int iv, int iv_to, int lvl, int add)
{
    for (int i = 0; i < n_ids(t, lvl); i++) {
        int id = level_id(t, lvl, i);
        prolong2(t, t->parent[id], id, iv, iv_to, add);
    }
    return 0;
}

/* a$D_prolong0 (m_aX_prolong.f90:39-108) over the child's interior */
static void prolong0(orc_tree *t, int p_id, int c_id, int iv, int ivc, int add)
{
    int n = t->ncell[c_id], np = t->ncell[p_id];
    int off[3];
    child_offset(t, c_id, off);
    const double *p = cc(t, p_id, iv);
    double *c = cc(t, c_id, ivc);
    int k0 = t->dim == 3 ? 1 : 0, k1 = t->dim == 3 ? n : 0;
    for (int k = k0; k <= k1; k++) {
        int kc = t->dim == 3 ? off[2] + ((k + 1) >> 1) : 0;
        for (int j = 1; j <= n; j++) {
            int jc = off[1] + ((j + 1) >> 1);
            for (int i = 1; i <= n; i++) {
                int ic = off[0] + ((i + 1) >> 1);
                double *q = &c[ix3(n, i, j, k)];
                if (add) *q = *q + p[ix3(np, ic, jc, kc)];
                else *q = p[ix3(np, ic, jc, kc)];
            }
        }
    }
}

/* a$D_prolong1 (m_aX_prolong.f90:130-219): 2-1-1 (2D) and 1-1-1-1 (3D)
 * linear interpolation, added to the old values or to a zeroed interior */
static void prolong1(orc_tree *t, int p_id, int c_id, int iv, int ivc, int add)
{
    int n = t->ncell[c_id], np = t->ncell[p_id], hnc = n >> 1;
    int off[3];
    child_offset(t, c_id, off);
    const double *p = cc(t, p_id, iv);
    double *c = cc(t, c_id, ivc);
    if (!add) {
        int kk = t->dim == 3 ? n : 0;
        for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
            for (int j = 1; j <= n; j++)
                for (int i = 1; i <= n; i++) c[ix3(n, i, j, k)] = 0;
    }
    if (t->dim == 2) {
        for (int j = 1; j <= hnc; j++) {
            int jc = j + off[1], jf = 2 * j - 1;
            for (int i = 1; i <= hnc; i++) {
                int ic = i + off[0], i_f = 2 * i - 1;
                double f0 = 0.5 * p[ix3(np, ic, jc, 0)];
                double flx = 0.25 * p[ix3(np, ic - 1, jc, 0)];
                double fhx = 0.25 * p[ix3(np, ic + 1, jc, 0)];
                double fly = 0.25 * p[ix3(np, ic, jc - 1, 0)];
                double fhy = 0.25 * p[ix3(np, ic, jc + 1, 0)];
                double *q;
                q = &c[ix3(n, i_f, jf, 0)];         *q = f0 + flx + fly + *q;
                q = &c[ix3(n, i_f + 1, jf, 0)];     *q = f0 + fhx + fly + *q;
                q = &c[ix3(n, i_f, jf + 1, 0)];     *q = f0 + flx + fhy + *q;
                q = &c[ix3(n, i_f + 1, jf + 1, 0)]; *q = f0 + fhx + fhy + *q;
            }
        }
        return;
    }
    for (int k = 1; k <= hnc; k++) {
        int kc = k + off[2], kf = 2 * k - 1;
        for (int j = 1; j <= hnc; j++) {
            int jc = j + off[1], jf = 2 * j - 1;
            for (int i = 1; i <= hnc; i++) {
                int ic = i + off[0], i_f = 2 * i - 1;
                double f0 = 0.25 * p[ix3(np, ic, jc, kc)];
                double fl[3] = {0.25 * p[ix3(np, ic - 1, jc, kc)], 0.25 * p[ix3(np, ic, jc - 1, kc)],
                                0.25 * p[ix3(np, ic, jc, kc - 1)]};
                double fh[3] = {0.25 * p[ix3(np, ic + 1, jc, kc)], 0.25 * p[ix3(np, ic, jc + 1, kc)],
                                0.25 * p[ix3(np, ic, jc, kc + 1)]};
                for (int s = 0; s < 8; s++) {
                    int sx = s & 1, sy = (s >> 1) & 1, sz = (s >> 2) & 1;
                    double *q = &c[ix3(n, i_f + sx, jf + sy, kf + sz)];
                    *q = f0 + (sx ? fh[0] : fl[0]) + (sy ? fh[1] : fl[1]) + (sz ? fh[2] : fl[2]) + *q;
                }
            }
        }
    }
}

/* a$D_prolong0/1/2_from (m_aX_prolong.f90:23-36, 112-125, 224-237) for every
 * parent of level lvl - 1, i.e. into every box of level lvl */
int orc_prolong_level(orc_tree *t, int iv, int iv_to, int lvl, int order, int add)
{
    if (order < 0 || order > 2 || lvl <= t->min_lvl) return 1;
    for (int i = 0; i < n_ids(t, lvl); i++) {
        int id = level_id(t, lvl, i);
        if (order == 0) prolong0(t, t->parent[id], id, iv, iv_to, add);
        else if (order == 1) prolong1(t, t->parent[id], id, iv, iv_to, add);
        else prolong2(t, t->parent[id], id, iv, iv_to, add);
    }
    return 0;
}

/* a$D_restrict_tree, m_aX_restrict.f90:47-57: the children of every parent,
 * highest level first */
int orc_restrict_tree(orc_tree *t, int iv, int iv_to)
{
    for (int lvl = t->highest_lvl; lvl > t->min_lvl; lvl--)
        for (int i = 0; i < n_ids(t, lvl); i++) {
            int id = level_id(t, lvl, i);
            restrict_box(t, id, t->parent[id], iv, iv_to);
        }
    return 0;
}

static void copy_all(orc_tree *t, int id, int from, int to)
{
    memcpy(cc(t, id, to), cc(t, id, from), (size_t)box_len(t, t->ncell[id]) * sizeof(double));
}

/* correct_children, m_aX_mg.f90:368-395, over the parents of level lvl-1 */
int orc_correct_children_level(orc_tree *t, const orc_mg_opts *o, int lvl)
{
    int pl = lvl - 1;
    for (int i = 0; i < n_ids(t, pl); i++) {
        int id = level_id(t, pl, i);
        if (!has_children(t, id)) continue;
        int len = box_len(t, t->ncell[id]);
        double *tm = cc(t, id, o->i_tmp);
        const double *ph = cc(t, id, o->i_phi);
        for (int q = 0; q < len; q++) tm[q] = ph[q] - tm[q];
        for (int c = 0; c < t->nch; c++) {
            int c_id = t->children[id * 8 + c];
            if (c_id == 0) continue;
            box_corr(t, id, c_id, o);
        }
    }
    return 0;
}

/* update_coarse, m_aX_mg.f90:424-480 */
int orc_update_coarse_level(orc_tree *t, const orc_mg_opts *o, int lvl)
{
    int nb = n_ids(t, lvl);
    for (int i = 0; i < nb; i++) {
        int id = level_id(t, lvl, i), p_id = t->parent[id];
        int n = t->ncell[id], len = box_len(t, n);
        double *tm = cc(t, id, o->i_tmp);
        double *save = malloc((size_t)len * sizeof(double));
        memcpy(save, tm, (size_t)len * sizeof(double));   /* interior is what matters */
        residual_box(t, id, o);
        box_rstr(t, id, p_id, o->i_tmp, o);
        box_rstr(t, id, p_id, o->i_phi, o);
        int kk = t->dim == 3 ? n : 0;
        for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
            for (int j = 1; j <= n; j++)
                for (int ii = 1; ii <= n; ii++) {
                    int q = ix3(n, ii, j, k);
                    tm[q] = save[q];
                }
        free(save);
    }
    orc_fill_gc_level(t, o->i_phi, lvl - 1);
    for (int i = 0; i < n_ids(t, lvl - 1); i++) {
        int id = level_id(t, lvl - 1, i);
        if (!has_children(t, id)) continue;
        box_op(t, id, o->i_rhs, o);
        int len = box_len(t, t->ncell[id]);
        double *r = cc(t, id, o->i_rhs), *tm = cc(t, id, o->i_tmp);
        for (int q = 0; q < len; q++) r[q] = r[q] + tm[q];   /* a$D_box_add_cc */
        copy_all(t, id, o->i_phi, o->i_tmp);                 /* a$D_box_copy_cc */
    }
    return 0;
}

/* set_coarse_phi_rhs, m_aX_mg.f90:484-517 */
static void set_coarse_phi_rhs(orc_tree *t, int lvl, const orc_mg_opts *o)
{
    if (lvl == t->highest_lvl) orc_fill_gc_level(t, o->i_phi, lvl);
    for (int i = 0; i < n_ids(t, lvl); i++) {
        int id = level_id(t, lvl, i), p_id = t->parent[id];
        residual_box(t, id, o);
        box_rstr(t, id, p_id, o->i_tmp, o);
        box_rstr(t, id, p_id, o->i_phi, o);
    }
    orc_fill_gc_level(t, o->i_phi, lvl - 1);
    for (int i = 0; i < n_ids(t, lvl - 1); i++) {
        int id = level_id(t, lvl - 1, i);
        if (!has_children(t, id)) continue;
        box_op(t, id, o->i_rhs, o);
        int len = box_len(t, t->ncell[id]);
        double *r = cc(t, id, o->i_rhs), *tm = cc(t, id, o->i_tmp);
        for (int q = 0; q < len; q++) r[q] = r[q] + tm[q];
    }
}

/* init_phi_rhs, m_aX_mg.f90:520-543 */
static void init_phi_rhs(orc_tree *t, const orc_mg_opts *o)
{
    for (int lvl = t->highest_lvl; lvl >= t->min_lvl + 1; lvl--)
        for (int i = 0; i < n_ids(t, lvl); i++) {
            int id = level_id(t, lvl, i);
            memset(cc(t, id, o->i_phi), 0, (size_t)box_len(t, t->ncell[id]) * sizeof(double));
            box_rstr(t, id, t->parent[id], o->i_rhs, o);
        }
}

static int check_opts(const orc_tree *t, const orc_mg_opts *o)
{
    if (o->i_phi < 1 || o->i_phi > t->n_var || o->i_rhs < 1 || o->i_rhs > t->n_var ||
        o->i_tmp < 1 || o->i_tmp > t->n_var)
        return -1;
    if (o->op == 1 && o->i_eps != -1 && (o->i_eps < 1 || o->i_eps > t->n_var)) return -1;
    if (o->op < 0 || o->op > 2 || o->rstr < 0 || o->rstr > 2) return -1;
    if ((o->i_lsf == -1) != (o->i_bval == -1)) return -1;   /* m_aX_mg.f90:136-137 */
    if (o->i_lsf != -1 && (o->i_lsf < 1 || o->i_lsf > t->n_var || o->i_bval < 1 || o->i_bval > t->n_var))
        return -1;
    if ((o->op == 2 || o->rstr == 2) && o->i_lsf == -1) return -1;
    if (o->n_cycle_down < 0 || o->n_cycle_up < 0 || o->n_cycle_base < 0) return -1;
    return 0;
}

/* mg$D_fas_vcycle, m_aX_mg.f90:208-255 */
int orc_fas_vcycle(orc_tree *t, const orc_mg_opts *o, int highest_lvl, int set_residual)
{
    if (check_opts(t, o)) return -1;
    int min_lvl = t->min_lvl;
    for (int lvl = highest_lvl; lvl >= min_lvl + 1; lvl--) {
        orc_gsrb_level(t, o, lvl, o->n_cycle_down);
        orc_update_coarse_level(t, o, lvl);
    }
    orc_gsrb_level(t, o, min_lvl, o->n_cycle_base);
    for (int lvl = min_lvl + 1; lvl <= highest_lvl; lvl++) {
        orc_correct_children_level(t, o, lvl);
        orc_fill_gc_level(t, o->i_phi, lvl);
        orc_gsrb_level(t, o, lvl, o->n_cycle_up);
    }
    if (set_residual)
        for (int lvl = min_lvl; lvl <= highest_lvl; lvl++) orc_residual_level(t, o, lvl);
    return t->tag_err ? -1 : 0;
}

/* mg$D_fas_fmg, m_aX_mg.f90:164-203 */
int orc_fas_fmg(orc_tree *t, const orc_mg_opts *o, int set_residual, int have_guess)
{
    if (check_opts(t, o)) return -1;
    int min_lvl = t->min_lvl;
    if (have_guess) {
        for (int lvl = t->highest_lvl; lvl >= min_lvl + 1; lvl--) set_coarse_phi_rhs(t, lvl, o);
    } else {
        init_phi_rhs(t, o);
    }
    for (int lvl = min_lvl; lvl <= t->highest_lvl; lvl++) {
        for (int i = 0; i < n_ids(t, lvl); i++)
            copy_all(t, level_id(t, lvl, i), o->i_phi, o->i_tmp);   /* a$D_boxes_copy_cc */
        if (lvl > min_lvl) {
            orc_correct_children_level(t, o, lvl);
            orc_fill_gc_level(t, o->i_phi, lvl);
        }
        orc_fas_vcycle(t, o, lvl, set_residual && lvl == t->highest_lvl);
    }
    return t->tag_err ? -1 : 0;
}

/* a$D_tree_max_cc / a$D_tree_min_cc, m_aX_utils.f90:468-571 */
static double tree_ext(const orc_tree *t, int iv, int include_parents, int want_max)
{
    double m = want_max ? -DBL_MAX : DBL_MAX;
    for (int lvl = t->min_lvl; lvl <= t->highest_lvl; lvl++)
        for (int i = 0; i < n_ids(t, lvl); i++) {
            int id = level_id(t, lvl, i);
            if (!include_parents && has_children(t, id)) continue;
            int n = t->ncell[id];
            const double *c = cc(t, id, iv);
            int kk = t->dim == 3 ? n : 0;
            for (int k = t->dim == 3 ? 1 : 0; k <= kk; k++)
                for (int j = 1; j <= n; j++)
                    for (int ii = 1; ii <= n; ii++) {
                        double v = c[ix3(n, ii, j, k)];
                        if (want_max ? v > m : v < m) m = v;
                    }
        }
    return m;
}

double orc_tree_max_cc(const orc_tree *t, int iv, int include_parents)
{
    return tree_ext(t, iv, include_parents, 1);
}

double orc_tree_min_cc(const orc_tree *t, int iv, int include_parents)
{
    return tree_ext(t, iv, include_parents, 0);
}

/* One box procedure of the family `kind` (0 lpl, 1 lpld, 2 lpllsf, 3 clpl,
 * 4 clpld) on one box, as a program calling it directly does:
 *   what 0: box op into i_out = arg (mg$D_box_lpl, _lpld, _lpllsf,
 *           mg2_box_clpl, _clpld)
 *   what 1: box gsrb with redblack_cntr = arg
 *   what 2: box corr of child id from its parent (kinds 0-2)
 *   what 3: box rstr of variable arg from child id into its parent
 *           (kind 0: a$D_restrict_box, by the tree's coordinates; 2: lpllsf)
 *   what 4: mg$D_sides_rb of face arg (1-based) of box id, variable o->i_phi */
int orc_box_call(orc_tree *t, const orc_mg_opts *o, int what, int kind, int id, int arg)
{
    if (id < 1 || id > t->max_id || t->ncell[id] == 0 || kind < 0 || kind > 4) return -1;
    if (kind >= 3 && !(t->dim == 2 && t->cyl)) return -1;
    int p = t->parent[id];
    switch (what) {
    case 0: {
        void (*f[5])(orc_tree *, int, int, const orc_mg_opts *) = {box_lpl, box_lpld, box_lpllsf, box_clpl,
                                                                     box_clpld};
        f[kind](t, id, arg, o);
        return 0;
    }
    case 1: {
        void (*f[5])(orc_tree *, int, int, const orc_mg_opts *) = {gsrb_lpl, gsrb_lpld, gsrb_lpllsf, gsrb_clpl,
                                                                     gsrb_clpld};
        f[kind](t, id, arg, o);
        return 0;
    }
    case 2:
        if (p <= 0 || kind > 2) return -1;
        if (kind == 0) prolong2(t, p, id, o->i_tmp, o->i_phi, 1);
        else if (kind == 1) corr_lpld(t, p, id, o);
        else corr_lpllsf(t, p, id, o);
        return 0;
    case 3:
        if (p <= 0 || (kind != 0 && kind != 2)) return -1;
        if (kind == 0) restrict_box(t, id, p, arg, arg);
        else rstr_lpllsf(t, id, p, arg, o);
        return 0;
    case 4:
        if (p <= 0 || arg < 1 || arg > 2 * t->dim) return -1;
        sides_rb(t, id, arg - 1, o->i_phi);
        return 0;
    }
    return -1;
}

/* a$D_gc2_box (m_aX_gc.f90:420-456) of variable iv on box id into
 * gc(nc[, nc], 2*dim) (Fortran order): sides2_from_nb (:459-495) from a
 * same-level neighbour, a$D_gc2_prolong1 (:498-597) at a refinement boundary,
 * a$D_bc2_neumann_zero (:600-634) at a physical boundary. */
static void gc2_box(orc_tree *t, int id, int iv, double *gc)
{
    int n = t->ncell[id], per = t->dim == 2 ? n : n * n;
    const double *own = cc(t, id, iv);
    for (int f = 0; f < t->nnb; f++) {
        int nb_id = t->nbr[id * 6 + f], dn = f / 2, high = f % 2;
        double *side = gc + (size_t)f * per;
        if (nb_id != 0) {
            /* sides2_from_nb: the neighbour's cell nc-1 (low face) or 2 (high
             * face); bc2_neumann_zero: the box's own cell 2 (low) or nc-1 (high) */
            const double *src = nb_id > 0 ? cc(t, nb_id, iv) : own;
            int m = (nb_id > 0) == (high != 0) ? 2 : n - 1;
            for (int b = 1; b <= (t->dim == 3 ? n : 1); b++)
                for (int a = 1; a <= n; a++) {
                    int x[3] = {0, 0, 0};
                    x[dn] = m;
                    x[dn == 0 ? 1 : 0] = a;
                    if (t->dim == 3) x[dn == 2 ? 1 : 2] = b;
                    side[(a - 1) + n * (b - 1)] = src[ix3(n, x[0], x[1], x[2])];
                }
            continue;
        }
        /* a$D_gc2_prolong1: ix_offset = a$D_get_child_offset(box, nb), the
         * offset minus dix(nb) * n_cell (lvl > 1) or n_cell / 2 (lvl <= 1) */
        int p = t->parent[id], pnb = t->nbr[p * 6 + f], np = t->ncell[pnb];
        int off[3];
        child_offset(t, id, off);
        off[dn] -= (high ? 1 : -1) * (t->lvl[id] > 1 ? n : n / 2);
        const double *q = cc(t, pnb, iv);
        int ix = high ? n + 2 : -1;
        int c1n = off[dn] + ((ix + 1) >> 1), c2n = c1n + 1 - 2 * (ix & 1);
        for (int b = 1; b <= (t->dim == 3 ? n : 1); b++)
            for (int a = 1; a <= n; a++) {
                int ta = dn == 0 ? 1 : 0, tb = dn == 2 ? 1 : 2;
                int c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0};
                c1[dn] = c1n;
                c2[dn] = c2n;
                c1[ta] = off[ta] + ((a + 1) >> 1);
                c2[ta] = c1[ta] + 1 - 2 * (a & 1);
                if (t->dim == 3) {
                    c1[tb] = off[tb] + ((b + 1) >> 1);
                    c2[tb] = c1[tb] + 1 - 2 * (b & 1);
                }
                /* terms in the order i, j(, k): cc(c1), cc(c2 in x), cc(c2 in y)(, cc(c2 in z)) */
                double v0 = q[ix3(np, c1[0], c1[1], c1[2])];
                double vx = q[ix3(np, c2[0], c1[1], c1[2])];
                double vy = q[ix3(np, c1[0], c2[1], c1[2])];
                double g;
                if (t->dim == 2)
                    g = 0.5 * v0 + 0.25 * vx + 0.25 * vy;
                else
                    g = 0.25 * v0 + 0.25 * vx + 0.25 * vy + 0.25 * q[ix3(np, c1[0], c1[1], c2[2])];
                side[(a - 1) + n * (b - 1)] = g;
            }
    }
}

int orc_gc2(orc_tree *t, int iv, double *out)
{
    size_t pos = 0;
    for (int r = 0; r < t->n_boxes; r++) {
        int id = t->rec_ids[r], n = t->ncell[id];
        gc2_box(t, id, iv, out + pos);
        pos += (size_t)2 * t->dim * (t->dim == 2 ? n : n * n);
    }
    return 0;
}
