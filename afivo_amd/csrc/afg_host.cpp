// afg_host.cpp -- the per-box operators of the multigrid module on the host
// (CPU): the public mg$D_box_* / mg$D_sides_rb routines of m_aX_mg for one box
// in host memory, for callers that apply them directly (the drop-in module's
// operator procedures).  They run the same per-cell bodies as the device
// kernels (afg_bodies.h, compiled here for the host), so a box gets the
// values the device stages give it.  Box arrays are box%cc(0:nc+1, ...) of
// one variable, x fastest.
#define AFG_DEV inline
#define AFG_HD
#include "afg_bodies.h"

#include <vector>

#include "afivo_amd.h"

using namespace afg;

namespace afg {
void set_call_error(const char *msg);   // the message afg_last_error(NULL) returns
}

namespace {

int fail(const char *msg)
{
    set_call_error(msg);
    return AFG_ERR_ARG;
}

int box_len_of(int dim, int nc) { return dim == 2 ? (nc + 2) * (nc + 2) : (nc + 2) * (nc + 2) * (nc + 2); }

bool bad_args(int dim, int nc) { return (dim != 2 && dim != 3) || nc < 2 || (nc & 1); }

// one box in slot 0 (and its parent in slot 1) of pools of box_len values
struct Scratch {
    int dim, nc, len;
    int32_t nb[12], parent[2], choff[2], bc_type[1] = {-1}, bc_voff[1] = {-1}, cyl_ioff[2];
    int8_t tag[2], is_parent[2] = {0, 1};
    double cyl_rmin[2], bc_vals[1] = {0};
    DevTree T;
    Scratch(int d, int n, int pnc) : dim(d), nc(n), len(box_len_of(d, n > pnc ? n : pnc))
    {
        for (int f = 0; f < 12; f++) nb[f] = NB_REFINE;
        parent[0] = 1;
        parent[1] = -1;
        choff[0] = choff[1] = 0;
        tag[0] = tag[1] = TAG_NORMAL;
        cyl_ioff[0] = cyl_ioff[1] = 0;
        cyl_rmin[0] = cyl_rmin[1] = 0;
        T.stride = len;
        T.nb = nb;
        T.parent = parent;
        T.choff = choff;
        T.bc_type = bc_type;
        T.bc_voff = bc_voff;
        T.bc_vals = bc_vals;
        T.tag = tag;
        T.is_parent = is_parent;
        T.cyl_ioff = nullptr;
        T.cyl_rmin = nullptr;
    }
    // a pool of two slots holding the child's (slot 0) and parent's (slot 1) arrays
    std::vector<double> pool(const double *child, const double *par, int pnc) const
    {
        std::vector<double> p(2 * (size_t)len, 0.0);
        if (child)
            for (int q = 0; q < box_len_of(dim, nc); q++) p[q] = child[q];
        if (par)
            for (int q = 0; q < box_len_of(dim, pnc); q++) p[len + q] = par[q];
        return p;
    }
};

LevelArgs level(int nc, int pnc, double dr)
{
    LevelArgs L;
    L.slot0 = 0;
    L.n = 1;
    L.nc = nc;
    L.pnc = pnc;
    L.tile = 0;
    L.dr = dr;
    L.dx2 = dr * dr;
    L.inv_dr_sq = 1 / (dr * dr);
    return L;
}

// the operator kinds of afg_box_op / afg_box_gsrb and the arrays they read
bool kind_ok(int dim, int nc, int kind, const double *eps, const double *lsf, const double *bval)
{
    if (bad_args(dim, nc) || kind < AFG_BOX_LPL || kind > AFG_BOX_CLPLD) return false;
    if ((kind == AFG_BOX_LPLD || kind == AFG_BOX_CLPLD) && !eps) return false;
    if (kind == AFG_BOX_LSF && (!lsf || !bval)) return false;
    return dim == 2 || kind < AFG_BOX_CLPL;
}

// a child with nc cells whose octant starts at offset (dim values) in a
// parent with pnc cells: pnc = nc (levels above 1) or nc / 2 (below, where
// the offset is 0)
bool child_ok(int dim, int nc, int pnc, const int32_t *offset)
{
    if (bad_args(dim, nc) || (pnc != nc && 2 * pnc != nc) || !offset) return false;
    for (int d = 0; d < dim; d++)
        if (offset[d] < 0 || offset[d] + nc / 2 > pnc) return false;
    return true;
}

int choff_of(const int32_t *off, int dim) { return off[0] | (off[1] << 10) | (dim == 3 ? off[2] << 20 : 0); }

}  // namespace

extern "C" {

int afg_box_op(int dim, int nc, double dr, int kind, int ir0, const double *phi, const double *eps,
               const double *lsf, const double *bval, double *out)
{
    if (!kind_ok(dim, nc, kind, eps, lsf, bval) || !phi || !out)
        return fail("afg_box_op: invalid dim, nc or kind, or an array the kind needs is missing");
    Scratch S(dim, nc, nc);
    const LevelArgs L = level(nc, nc, dr);
    const LevelView v{const_cast<double *>(phi), nullptr, nullptr, eps, 0, S.len};
    const int s = nc + 2;
    for (int k = dim == 3 ? 1 : 0; k <= (dim == 3 ? nc : 0); k++)
        for (int j = 1; j <= nc; j++)
            for (int i = 1; i <= nc; i++) {
                const int q = i + s * (j + s * k);
                double u[6];
                double a;
                if (dim == 2) {
                    stencil<2>(S.T, L, v, v, 0, phi, nc, i, j, k, q, 0, phi[q], u);
                    a = kind == AFG_BOX_LSF ? box_op_lsf<2>(u, phi[q], lsf, bval, q, s, L.inv_dr_sq)
                        : kind >= AFG_BOX_CLPL
                            ? box_op_cyl(u, phi[q], eps, q, s, L.inv_dr_sq, kind == AFG_BOX_CLPLD, ir0 + i)
                            : box_op<2>(u, phi[q], eps, q, s, L.inv_dr_sq, kind == AFG_BOX_LPLD);
                } else {
                    stencil<3>(S.T, L, v, v, 0, phi, nc, i, j, k, q, 0, phi[q], u);
                    a = kind == AFG_BOX_LSF ? box_op_lsf<3>(u, phi[q], lsf, bval, q, s, L.inv_dr_sq)
                                            : box_op<3>(u, phi[q], eps, q, s, L.inv_dr_sq, kind == AFG_BOX_LPLD);
                }
                out[q] = a;
            }
    return 0;
}

int afg_box_gsrb(int dim, int nc, double dr, int kind, int ir0, int redblack, double *phi, const double *rhs,
                 const double *eps, const double *lsf, const double *bval)
{
    if (!kind_ok(dim, nc, kind, eps, lsf, bval) || !phi || !rhs)
        return fail("afg_box_gsrb: invalid dim, nc or kind, or an array the kind needs is missing");
    Scratch S(dim, nc, nc);
    const LevelArgs L = level(nc, nc, dr);
    const LevelView v{phi, nullptr, nullptr, eps, 0, S.len};
    const int s = nc + 2;
    const int half = (dim == 2 ? nc * nc : nc * nc * nc) / 2;
    // one colour: every cell reads only cells of the other colour and ghosts
    for (int c = 0; c < half; c++) {
        int i, j, k;
        double u[6];
        if (dim == 2) {
            colour_cell<2>(c, nc, redblack, i, j, k);
            const int q = i + s * j;
            stencil<2>(S.T, L, v, v, 0, phi, nc, i, j, k, q, 0, phi[q], u);
            phi[q] = kind == AFG_BOX_LSF ? gsrb_value_lsf<2>(u, rhs[q], lsf, bval, q, s, L.dx2)
                     : kind >= AFG_BOX_CLPL
                         ? gsrb_value_cyl(u, rhs[q], eps, q, s, L.dx2, kind == AFG_BOX_CLPLD, ir0 + i)
                         : gsrb_value<2>(u, rhs[q], eps, q, s, L.dx2, kind == AFG_BOX_LPLD);
        } else {
            colour_cell<3>(c, nc, redblack, i, j, k);
            const int q = i + s * (j + s * k);
            stencil<3>(S.T, L, v, v, 0, phi, nc, i, j, k, q, 0, phi[q], u);
            phi[q] = kind == AFG_BOX_LSF ? gsrb_value_lsf<3>(u, rhs[q], lsf, bval, q, s, L.dx2)
                                         : gsrb_value<3>(u, rhs[q], eps, q, s, L.dx2, kind == AFG_BOX_LPLD);
        }
    }
    return 0;
}

int afg_box_corr(int dim, int nc, int pnc, int kind, const int32_t *offset, const double *p_corr,
                 const double *p_eps, const double *p_lsf, double *c_phi, const double *c_lsf)
{
    const bool need = kind == AFG_BOX_LPL || (kind == AFG_BOX_LPLD && p_eps) || (kind == AFG_BOX_LSF && p_lsf && c_lsf);
    if (!child_ok(dim, nc, pnc, offset) || !need || !p_corr || !c_phi)
        return fail("afg_box_corr: invalid dim, nc, pnc, offset or kind, or an array the kind needs is missing");
    Scratch S(dim, nc, pnc);
    S.choff[0] = choff_of(offset, dim);
    if (kind == AFG_BOX_LPLD) S.tag[0] = TAG_VEPS;   // the lpld correction for this box
    const LevelArgs L = level(nc, pnc, 1.0);
    std::vector<double> phi = S.pool(c_phi, nullptr, pnc), tmp = S.pool(nullptr, p_corr, pnc);
    std::vector<double> eps = S.pool(nullptr, p_eps, pnc), lsf = S.pool(c_lsf, p_lsf, pnc);
    LevelView v{phi.data(), nullptr, tmp.data(), eps.data(), 0, S.len, lsf.data(), nullptr,
                kind == AFG_BOX_LSF ? 2 : 0, 0};
    const int per = dim == 2 ? (nc / 2) * (nc / 2) : (nc / 2) * (nc / 2) * (nc / 2);
    for (int g = 0; g < per; g++) {
        if (dim == 2)
            correct_body<2, 0>(S.T, L, v, v, kind == AFG_BOX_LPLD, 0, g);
        else
            correct_body<3, 0>(S.T, L, v, v, kind == AFG_BOX_LPLD, 0, g);
    }
    for (int q = 0; q < box_len_of(dim, nc); q++) c_phi[q] = phi[q];
    return 0;
}

int afg_box_rstr(int dim, int nc, int pnc, int kind, const int32_t *offset, const double *c_var,
                 const double *c_lsf, const double *p_lsf, double p_rmin, double c_dr, double *p_var)
{
    const bool need = kind == AFG_BOX_LPL || (kind == AFG_BOX_CLPL && dim == 2 && p_rmin >= 0 && c_dr > 0) ||
                      (kind == AFG_BOX_LSF && p_lsf && c_lsf);
    if (!child_ok(dim, nc, pnc, offset) || !need || !c_var || !p_var)
        return fail("afg_box_rstr: invalid dim, nc, pnc, offset or kind, or an array the kind needs is missing");
    Scratch S(dim, nc, pnc);
    S.choff[0] = choff_of(offset, dim);
    if (kind == AFG_BOX_CLPL) {
        S.T.cyl_rmin = S.cyl_rmin;
        S.T.cyl_ioff = S.cyl_ioff;
        S.cyl_rmin[1] = p_rmin;
    }
    const LevelArgs L = level(nc, pnc, c_dr);
    std::vector<double> var = S.pool(c_var, p_var, pnc), lsf = S.pool(c_lsf, p_lsf, pnc);
    const int per = dim == 2 ? (nc / 2) * (nc / 2) : (nc / 2) * (nc / 2) * (nc / 2);
    for (int g = 0; g < per; g++) {
        if (dim == 2)
            restrict_body<2, 0>(S.T, L, var.data(), var.data(), g, lsf.data(), kind == AFG_BOX_LSF ? 2 : 0);
        else
            restrict_body<3, 0>(S.T, L, var.data(), var.data(), g, lsf.data(), kind == AFG_BOX_LSF ? 2 : 0);
    }
    for (int q = 0; q < box_len_of(dim, pnc); q++) p_var[q] = var[S.len + q];
    return 0;
}

int afg_box_sides_rb(int dim, int nc, int pnc, int nb, const int32_t *offset, const double *p_var, double *c_var)
{
    if (!child_ok(dim, nc, pnc, offset) || nb < 1 || nb > 2 * dim || !p_var || !c_var)
        return fail("afg_box_sides_rb: invalid dim, nc, pnc, offset or face, or a missing array");
    Scratch S(dim, nc, pnc);
    S.choff[0] = choff_of(offset, dim);
    const LevelArgs L = level(nc, pnc, 1.0);
    std::vector<double> var = S.pool(c_var, p_var, pnc);
    const LevelView v{var.data(), nullptr, nullptr, nullptr, 0, S.len};
    const int per_face = dim == 2 ? nc : nc * nc;
    for (int t = 0; t < per_face; t++) {
        const int g = (nb - 1) * per_face + t;   // the face's threads of fill_body
        if (dim == 2)
            fill_body<2, 0>(S.T, L, v, v, g);
        else
            fill_body<3, 0>(S.T, L, v, v, g);
    }
    for (int q = 0; q < box_len_of(dim, nc); q++) c_var[q] = var[q];
    return 0;
}

}  // extern "C"
