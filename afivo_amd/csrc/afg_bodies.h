// afg_bodies.h -- per-thread bodies of the afivo_amd multigrid stages.
//
// Every stage restates one afivo routine over all boxes of a level (the
// reference runs these as OpenMP loops over boxes).  Floating-point
// expressions keep the reference's evaluation order and the library is built
// with -ffp-contract=off, so results are bit-identical to the reference.
//
// A body computes the work of one thread (index gid) of a stage; a stage is
// race-free, so its threads may run in any order.  Bodies address boxes
// through LevelView, so the same code runs on the device pools in HBM (one
// kernel launch per stage for the large levels) and on copies of the small
// levels held in LDS (the bottom kernel, one launch for the whole lower part
// of a V-cycle).  Box arrays hold cc(0:nc+1, 0:nc+1[, 0:nc+1]) with x fastest;
// s = nc + 2 is the row pitch.  NC is the compile-time cell count per
// dimension (0: taken from the level arguments).
//
// AFG_DEV / AFG_HD default to the HIP device qualifiers; a host build of the
// bodies (used to check the stage logic on the CPU) defines them first.
#pragma once

#include <type_traits>

#include "afg_types.h"

#ifndef AFG_DEV
#define AFG_DEV __device__ __forceinline__
#endif
#ifndef AFG_HD
#define AFG_HD __host__ __device__
#endif

namespace afg {

AFG_HD constexpr int ipow(int n, int d) { return d == 2 ? n * n : n * n * n; }

// Loads of the multigrid variables (phi, rhs, tmp: the arrays that stages
// write).  Where the workgroups of one launch exchange these arrays between
// its steps (the persistent kernel, compiled with AFG_COHERENT_LOADS = 1),
// they are read with device-scope loads: those miss in the compute unit's
// vector L1, which is not kept coherent with the stores of other compute
// units, and read the L2 that the workgroups of one XCD share.  Elsewhere
// they are plain loads.  Tables, eps and the level set are never written
// within a launch and are read directly.
#ifndef AFG_COHERENT_LOADS
#define AFG_COHERENT_LOADS 0
#endif
template <typename R>
AFG_DEV R ldv(const R *p)
{
#if AFG_COHERENT_LOADS
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    return *p;
#endif
}

// Where the boxes of one level live: box `slot` of variable X starts at
// X + (slot - slot0) * stride.  The HBM pools use slot0 = 0 and the pool
// stride for every level; an LDS copy of a level uses its own slot0 and the
// level's box size.
// R is the value type: double for the reference arithmetic, float for the
// FP32 stages of the mixed-precision solve.
template <typename R>
struct LevelViewT {
    R *phi, *rhs, *tmp;
    const R *eps;
    int slot0, stride;
    // level set and boundary value of the lpllsf operators, with the modes of MgVarsT
    const R *lsf = nullptr, *bval = nullptr;
    int lsf_mode = 0, rstr_mode = 0;
    // rhs by colour (MgVarsT::rhs_rb), or nullptr
    R *rhs_rb = nullptr;
    int rb_stride = 0;
    AFG_DEV size_t off(int slot) const
    {
        return (size_t)(slot - slot0) * (size_t)stride;
    }
};
using LevelView = LevelViewT<double>;

// ---------------------------------------------------------------------------
// ghost value of face f next to interior cell (i, j, k) of box `slot`
//
// The one expression used by the ghost fills and, on the fly, by the stages
// that read face neighbours.  Restates a$D_gc_box (m_aX_gc.f90:50-69):
//   neighbour at the same level  -> sides_from_nb  (m_aX_gc.f90:381-415)
//   refinement boundary          -> mg$D_sides_rb  (m_aX_mg.f90:262-350) after
//                                   a$D_gc_prolong0 (m_aX_gc.f90:135-151,
//                                   m_aX_prolong.f90:39-108)
//   physical boundary            -> captured sides_bc value b, then bc_to_gc
//                                   (m_aX_gc.f90:71-132): c2*b + c1*interior
// `center` is phi at (i, j, k); `own` holds the level, `par` its parent level;
// `xbox`, when given, holds the box's own cells in place of own's (same
// indexing), for a box whose current values are in LDS.
template <int DIM, typename R>
AFG_DEV R ghost_value(const DevTree &T, const LevelViewT<R> &own, const LevelViewT<R> &par, int pnc,
                      double dr, int slot, int f, int nc, int i, int j, int k, R center,
                      const R *xbox = nullptr)
{
    const int s = nc + 2;
    const int nbv = T.nb[slot * (2 * DIM) + f];
    const int dn = f >> 1, high = f & 1;
    if (nbv >= 0) {
        const int m = high ? 1 : nc;
        const int a0 = dn == 0 ? m : i, a1 = dn == 1 ? m : j, a2 = dn == 2 ? m : k;
        return ldv(own.phi + own.off(nbv) + a0 + s * (a1 + s * a2));
    }
    if (nbv == NB_REFINE) {
        const int p = T.parent[slot];
        const int co = T.choff[slot];
        const int gv = high ? nc + 1 : 0;
        const int g0 = dn == 0 ? gv : i, g1 = dn == 1 ? gv : j, g2 = dn == 2 ? gv : k;
        const int ps = pnc + 2;
        const int pi = (co & 1023) + ((g0 + 1) >> 1);
        const int pj = ((co >> 10) & 1023) + ((g1 + 1) >> 1);
        const int pk = DIM == 3 ? ((co >> 20) & 1023) + ((g2 + 1) >> 1) : 0;
        const R P = ldv(par.phi + par.off(p) + pi + ps * (pj + ps * pk));
        const R *box = xbox ? xbox : own.phi + own.off(slot);
        const int inward = high ? -1 : 1;
        if (DIM == 2) {
            const int di = dn == 0 ? inward : -1 + 2 * (i & 1);
            const int dj = dn == 1 ? inward : -1 + 2 * (j & 1);
            const R a = ldv(box + (i + di) + s * j);
            const R b = ldv(box + i + s * (j + dj));
            return R(0.5) * P + center - R(0.25) * (a + b);
        } else {
            const int d0 = dn == 0 ? inward : -1 + 2 * (i & 1);
            const int d1 = dn == 1 ? inward : -1 + 2 * (j & 1);
            const int d2 = dn == 2 ? inward : -1 + 2 * (k & 1);
            const R d = ldv(box + (i + d0) + s * ((j + d1) + s * (k + d2)));
            return R(0.5) * P + R(0.75) * center - R(0.25) * d;
        }
    }
    const int rec = NB_BC0 - nbv;
    const int voff = T.bc_voff[rec];
    R b = 0;
    if (voff >= 0) {
        const int t1 = dn == 0 ? j : i;
        const int t2 = dn == 2 ? j : k;
        b = R(T.bc_vals[voff + (t1 - 1) + (DIM == 3 ? nc * (t2 - 1) : 0)]);
    }
    R c1, c2;
    if (T.bc_type[rec] == -1) {   // a5_bc_dirichlet
        c1 = -1;
        c2 = 2;
    } else {                      // a5_bc_neumann
        c1 = 1;
        c2 = R(dr * (high ? 1.0 : -1.0));
    }
    return c2 * b + c1 * center;
}

// The 2*DIM face neighbours of interior cell q = (i, j, k), in the order
// i-1, i+1, j-1, j+1, k-1, k+1.  A neighbour in the ghost layer is read from
// memory when `fly` is 0 and computed with ghost_value when `fly` is 1.
template <int DIM, typename R>
AFG_DEV void stencil(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &own,
                     const LevelViewT<R> &par, int slot, const R *box, int nc, int i, int j, int k,
                     int q, int fly, R center, R u[6])
{
    const int s = nc + 2;
    u[0] = (i > 1 || !fly) ? ldv(box + q - 1) : ghost_value<DIM>(T, own, par, L.pnc, L.dr, slot, 0, nc, i, j, k, center);
    u[1] = (i < nc || !fly) ? ldv(box + q + 1) : ghost_value<DIM>(T, own, par, L.pnc, L.dr, slot, 1, nc, i, j, k, center);
    u[2] = (j > 1 || !fly) ? ldv(box + q - s) : ghost_value<DIM>(T, own, par, L.pnc, L.dr, slot, 2, nc, i, j, k, center);
    u[3] = (j < nc || !fly) ? ldv(box + q + s) : ghost_value<DIM>(T, own, par, L.pnc, L.dr, slot, 3, nc, i, j, k, center);
    if (DIM == 3) {
        u[4] = (k > 1 || !fly) ? ldv(box + q - s * s) : ghost_value<DIM>(T, own, par, L.pnc, L.dr, slot, 4, nc, i, j, k, center);
        u[5] = (k < nc || !fly) ? ldv(box + q + s * s) : ghost_value<DIM>(T, own, par, L.pnc, L.dr, slot, 5, nc, i, j, k, center);
    } else {
        u[4] = u[5] = 0;
    }
}

// A phi at interior cell q from its neighbours u: mg$D_box_lpl
// (m_aX_mg.f90:763-797) or, for tagged boxes, mg$D_box_lpld (:866-916).
// E: the type eps is stored in (R, or float for FP64 arithmetic on FP32 values).
template <int DIM, typename R, typename E = R>
AFG_DEV R box_op(const R u[6], R u0, const E *e, int q, int s, double inv_dr_sq, bool lpld)
{
    if (lpld) {
        const int dq[6] = {-1, 1, -s, s, -s * s, s * s};
        const R a0 = e[q];
        R sum = 0;
#pragma unroll
        for (int m = 0; m < 2 * DIM; m++) {
            const R a = e[q + dq[m]];
            const R t = a0 * a / (a0 + a) * (u[m] - u0);
            sum = m == 0 ? t : sum + t;
        }
        return R(inv_dr_sq) * 2 * sum;
    }
    if (DIM == 2) return R(inv_dr_sq) * (u[0] + u[1] + u[2] + u[3] - 4 * u0);
    return R(inv_dr_sq) * (u[0] + u[1] + u[2] + u[3] + u[4] + u[5] - 6 * u0);
}

AFG_DEV bool lpld_box(const DevTree &T, int op_auto, int slot)
{
    return op_auto && (T.tag[slot] == TAG_CEPS || T.tag[slot] == TAG_VEPS);
}

// ---- internal boundaries given by a level set (m_aX_mg.f90:987-1321) ----

// the lpllsf operators for this box: every box (the lpllsf set), or the
// boxes tagged mg_lsf_box (the auto set)
template <typename R>
AFG_DEV bool lsf_box(const DevTree &T, const LevelViewT<R> &v, int slot)
{
    return v.lsf_mode == 2 || (v.lsf_mode == 1 && T.tag[slot] == TAG_LSF);
}

// mg$D_box_rstr_lpllsf for this child: every box, or by its tag (mg$D_auto_rstr)
template <typename R>
AFG_DEV bool lsf_rstr_box(const DevTree &T, const LevelViewT<R> &v, int slot)
{
    return v.rstr_mode == 2 || (v.rstr_mode == 1 && T.tag[slot] == TAG_LSF);
}

// lsf_dist_val (m_aX_mg.f90:990-1017): distance in (0, 1] from a to the
// neighbour b and the value there; where the level set changes sign between
// them, the distance to the zero crossing and the boundary value
// interpolated to it
template <typename R>
AFG_DEV void lsf_dist_val(R la, R ba, R lb, R vb, R bb, R &dist, R &val)
{
    if (la * lb < 0) {
        const R d = la / (la - lb);
        dist = d;
        val = ba * (1 - d) + bb * d;
    } else {
        dist = 1;
        val = vb;
    }
}

// distances dd[1..2*DIM] and values val[1..] of the face neighbours u of
// cell q (stencil order i-1, i+1, j-1, j+1, k-1, k+1); l, b: the box's level
// set and boundary value
template <int DIM, typename R>
AFG_DEV void lsf_neighbours(const R u[6], const R *l, const R *b, int q, int s, R dd[7], R val[7])
{
    const int dq[6] = {-1, 1, -s, s, -s * s, s * s};
#pragma unroll
    for (int m = 0; m < 2 * DIM; m++)
        lsf_dist_val(l[q], b[q], l[q + dq[m]], u[m], b[q + dq[m]], dd[m + 1], val[m + 1]);
}

// mg$D_box_gsrb_lpllsf (m_aX_mg.f90:1105-1179): new phi at cell q
template <int DIM, typename R>
AFG_DEV R gsrb_value_lsf(const R u[6], R rhs, const R *l, const R *b, int q, int s, double dx2)
{
    R d[7], v[7];
    lsf_neighbours<DIM>(u, l, b, q, s, d, v);
    if (DIM == 2) {
        const R prod = d[1] * d[2] * d[3] * d[4];
        return R(1) / (d[1] * d[2] + d[3] * d[4]) *
               ((d[2] * v[1] + d[1] * v[2]) * d[3] * d[4] / (d[1] + d[2]) +
                (d[4] * v[3] + d[3] * v[4]) * d[1] * d[2] / (d[3] + d[4]) - R(0.5) * prod * R(dx2) * rhs);
    }
    return R(1) / (R(1) / (d[1] * d[2]) + R(1) / (d[3] * d[4]) + R(1) / (d[5] * d[6])) *
           ((d[2] * v[1] + d[1] * v[2]) / ((d[1] + d[2]) * d[1] * d[2]) +
            (d[4] * v[3] + d[3] * v[4]) / ((d[3] + d[4]) * d[3] * d[4]) +
            (d[6] * v[5] + d[5] * v[6]) / ((d[5] + d[6]) * d[5] * d[6]) - R(0.5) * R(dx2) * rhs);
}

// mg$D_box_lpllsf (m_aX_mg.f90:1182-1251): the operator at cell q with value f0
template <int DIM, typename R>
AFG_DEV R box_op_lsf(const R u[6], R f0, const R *l, const R *b, int q, int s, double inv_dr_sq)
{
    R d[7], v[7];
    lsf_neighbours<DIM>(u, l, b, q, s, d, v);
    R sum = (d[2] * v[1] + d[1] * v[2] - (d[1] + d[2]) * f0) / ((d[1] + d[2]) * d[1] * d[2]) +
            (d[4] * v[3] + d[3] * v[4] - (d[3] + d[4]) * f0) / ((d[3] + d[4]) * d[3] * d[4]);
    if (DIM == 3)
        sum = sum + (d[6] * v[5] + d[5] * v[6] - (d[5] + d[6]) * f0) / ((d[5] + d[6]) * d[5] * d[6]);
    return R(2) * R(inv_dr_sq) * sum;
}

// mg$D_box_rstr_lpllsf (m_aX_mg.f90:1254-1320) of one coarse cell: values v
// and level set lc of its 2^DIM children in Fortran order, level set lp of
// the parent cell.  Children on the parent cell's side of the interface
// only, where some but not all are; 1 / n_ch is integer division there, so
// 2..2^DIM-1 such children give 0 * sum.
template <int DIM, typename R>
AFG_DEV R restrict_lsf(const R v[8], const R lc[8], R lp)
{
    constexpr int NCH = 1 << DIM;
    bool mk[8];
    int n_ch = 0;
#pragma unroll
    for (int m = 0; m < NCH; m++) {
        mk[m] = lp * lc[m] > 0;
        n_ch += mk[m];
    }
    if (n_ch < NCH && n_ch > 0) {
        R sm = 0;
#pragma unroll
        for (int m = 0; m < NCH; m++)
            if (mk[m]) sm += v[m];
        return R(1 / n_ch) * sm;
    }
    R sm = v[0];
#pragma unroll
    for (int m = 1; m < NCH; m++) sm += v[m];
    return R(DIM == 2 ? 0.25 : 0.125) * sm;
}

// ---- cylindrical coordinates (2D, x the radius): the auto operators of a
// cylindrical tree, mg2_box_gsrb_clpl / mg2_box_clpl for untagged boxes and
// mg2_box_gsrb_clpld / mg2_box_clpld for eps boxes (m_aX_mg.f90:1323-1443),
// with the radial factors rfac = [ir - 1, ir] / (ir - 0.5) of global radial
// cell index ir = i + (ix(1) - 1) nc (1 for the y neighbours), and the
// radius-weighted restriction (m_aX_restrict.f90:90-103).
AFG_DEV bool cyl_op(const DevTree &T, int op_auto) { return T.cyl_ioff != nullptr && op_auto; }

template <typename R>
AFG_DEV void cyl_rfac(int ir, R rf[4])
{
    const double d = double(ir) - 0.5;
    rf[0] = R(double(ir - 1) / d);
    rf[1] = R(double(ir) / d);
    rf[2] = rf[3] = R(1);
}

// u in the order i-1, i+1, j-1, j+1
template <typename R>
AFG_DEV R gsrb_value_cyl(const R u[6], R rhs, const R *e, int q, int s, double dx2, bool lpld, int ir)
{
    R rf[4];
    cyl_rfac(ir, rf);
    if (lpld) {
        // c = 2 a0 a / (a0 + a); phi = (sum(c rfac u) - dx2 rhs) / sum(c rfac)
        const R a0 = e[q];
        const int dq[4] = {-1, 1, -s, s};
        R su = 0, sc = 0;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const R a = e[q + dq[m]];
            const R cm = 2 * a0 * a / (a0 + a);
            const R cr = m < 2 ? cm * rf[m] : cm;
            su = m == 0 ? cr * u[m] : su + cr * u[m];
            sc = m == 0 ? cr : sc + cr;
        }
        return (su - R(dx2) * rhs) / sc;
    }
    // order of m_aX_mg.f90:1345-1349: i-1, i+1, j+1, j-1
    return R(0.25) * (rf[0] * u[0] + rf[1] * u[1] + u[3] + u[2] - R(dx2) * rhs);
}

template <typename R>
AFG_DEV R box_op_cyl(const R u[6], R u0, const R *e, int q, int s, double inv_dr_sq, bool lpld, int ir)
{
    R rf[4];
    cyl_rfac(ir, rf);
    if (lpld) {
        const int dq[4] = {-1, 1, -s, s};
        const R a0 = e[q];
        R sum = 0;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const R a = e[q + dq[m]];
            const R t = (m < 2 ? rf[m] * a0 : a0) * a / (a0 + a) * (u[m] - u0);
            sum = m == 0 ? t : sum + t;
        }
        return R(inv_dr_sq) * 2 * sum;
    }
    return (rf[0] * u[0] + rf[1] * u[1] + u[2] + u[3] - 4 * u0) * R(inv_dr_sq);
}

// The weighted restriction of the children values v (in the order of the
// reference's sum: (i_f, j_f), (i_f+1, j_f), (i_f, j_f+1), (i_f+1, j_f+1)):
// 0.25 (w1 (v0 + v2) + w2 (v1 + v3)), w = 1 -/+ 0.25 dr_p / r with r the
// radius r_min(1) + (ii - 0.5) dr_p of the parent's cell ii -- the octant
// column, as a2_cyl_child_weights(box_p, i, ...) is called with the loop
// index (m_aX_restrict.f90:98, m_aX_t.f90:507-528).
template <typename R>
AFG_DEV R cyl_restrict(const R v[4], double rmin_p, double dr_p, int ii)
{
    const double r = rmin_p + (ii - 0.5) * dr_p;
    const double t = 0.25 * dr_p / r;
    const R w1 = R(1 - t), w2 = R(1 + t);
    return R(0.25) * (w1 * (v[0] + v[2]) + w2 * (v[1] + v[3]));
}

// ---------------------------------------------------------------------------
// ghost fill of a level: fill_gc_phi (m_aX_mg.f90:352-365)
// one thread per face ghost cell: n * 2*DIM * nc^(DIM-1)

template <int DIM>
AFG_HD inline int fill_threads(int n, int nc)
{
    return n * 2 * DIM * (DIM == 2 ? nc : nc * nc);
}

template <int DIM, int NC, typename R>
AFG_DEV void fill_body(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &own,
                       const LevelViewT<R> &par, int gid)
{
    const int nc = NC ? NC : L.nc;
    const int per_face = DIM == 2 ? nc : nc * nc;
    const int per_box = 2 * DIM * per_face;
    const int b = gid / per_box;
    const int r = gid - b * per_box;
    const int f = r / per_face;
    const int t = r - f * per_face;
    const int t1 = t % nc + 1, t2 = DIM == 3 ? t / nc + 1 : 0;
    const int dn = f >> 1, high = f & 1;
    const int in = high ? nc : 1;
    const int i = dn == 0 ? in : t1;
    const int j = dn == 1 ? in : (dn == 0 ? t1 : (DIM == 3 ? t2 : 0));
    const int k = DIM == 3 ? (dn == 2 ? in : t2) : 0;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    R *box = own.phi + own.off(slot);
    // the cell next to the ghost enters only at refinement and physical boundaries
    const R center = T.nb[slot * (2 * DIM) + f] >= 0 ? R(0) : ldv(box + i + s * (j + s * k));
    const R g = ghost_value<DIM>(T, own, par, L.pnc, L.dr, slot, f, nc, i, j, k, center);
    const int gi = dn == 0 ? (high ? nc + 1 : 0) : i;
    const int gj = dn == 1 ? (high ? nc + 1 : 0) : j;
    const int gk = dn == 2 ? (high ? nc + 1 : 0) : k;
    box[gi + s * (gj + s * gk)] = g;
}

// correct_children, parent part (m_aX_mg.f90:380-381): tmp = phi - tmp over
// the whole array of each parent.
template <int DIM, int NC, typename R>
AFG_DEV void diff_body(const DevTree &T, const ListArgs &P, const LevelViewT<R> &v, int gid)
{
    const int nc = NC ? NC : P.nc;
    const int per_box = ipow(nc + 2, DIM);
    const int b = gid / per_box;
    const int q = gid - b * per_box;
    const size_t o = v.off(P.slots[b]) + q;
    v.tmp[o] = ldv(v.phi + o) - ldv(v.tmp + o);
}

// ---------------------------------------------------------------------------
// one colour pass of GSRB over a level: the body of gsrb_boxes
// (m_aX_mg.f90:404-411) with mg$D_box_gsrb_lpl (m_aX_mg.f90:718-760) or, for
// tagged boxes under the auto operators, mg$D_box_gsrb_lpld (:810-863).
//
// With fly = 0 the face neighbours come from the ghost layer in memory, as in
// the reference.  With fly = 1 they are computed with ghost_value from the
// state at the start of the pass, which equals what the reference's ghost
// fill after the previous pass stores: a face ghost feeds only the interior
// cell next to it, and everything it depends on (that cell, the neighbour's
// mirrored cell, the parent level) is untouched by the pass, nc being even.
// One thread per cell of colour rb.

template <int DIM>
AFG_HD inline int gsrb_threads(int n, int nc)
{
    return n * ipow(nc, DIM) / 2;
}

// New phi at interior cell q of a box from its face neighbours u (order of
// stencil) and rhs: mg$D_box_gsrb_lpl (m_aX_mg.f90:718-760) or, when lpld,
// mg$D_box_gsrb_lpld (:810-863) with eps e of the box.
template <int DIM, typename R>
AFG_DEV R gsrb_value(const R u[6], R rhs, const R *e, int q, int s, double dx2, bool lpld)
{
    if (lpld) {
        // c = 2 a0 a / (a0 + a); phi = (sum(c u) - dx2 rhs) / sum(c)
        const R a0 = e[q];
        const int dq[6] = {-1, 1, -s, s, -s * s, s * s};
        R su = 0, sc = 0;
#pragma unroll
        for (int m = 0; m < 2 * DIM; m++) {
            const R a = e[q + dq[m]];
            const R cm = 2 * a0 * a / (a0 + a);
            su = m == 0 ? cm * u[m] : su + cm * u[m];
            sc = m == 0 ? cm : sc + cm;
        }
        return (su - R(dx2) * rhs) / sc;
    }
    if (DIM == 2) {
        // order of m_aX_mg.f90:740-743: i+1, i-1, j+1, j-1
        return R(0.25) * (u[1] + u[0] + u[3] + u[2] - R(dx2) * rhs);
    }
    constexpr double sixth = 1 / 6.0;
    return R(sixth) * (u[1] + u[0] + u[3] + u[2] + u[5] + u[4] - R(dx2) * rhs);
}

// interior cell (i, j, k) of colour rb for index c in [0, nc^DIM / 2): the
// cells with i + j (+ k) = rb (mod 2), i fastest
template <int DIM>
AFG_HD inline void colour_cell(int c, int nc, int rb, int &i, int &j, int &k)
{
    const int hn = nc >> 1;
    const int r = c / hn;
    const int p = c - r * hn;
    j = DIM == 2 ? r + 1 : r % nc + 1;
    k = DIM == 2 ? 0 : r / nc + 1;
    i = 2 - ((rb ^ (DIM == 2 ? j : k + j)) & 1) + 2 * p;
}

// position of interior cell (i, j, k) in the colour-split rhs of its box:
// colour rb - 1 block (rb = 1 where i + j + k is odd), index c of colour_cell
template <int DIM>
AFG_HD inline int colour_index(int i, int j, int k, int nc)
{
    const int hn = nc >> 1;
    const int r = DIM == 2 ? j - 1 : (j - 1) + nc * (k - 1);
    const int odd = (i + j + k) & 1;
    return (odd ? 0 : ipow(nc, DIM) / 2) + r * hn + ((i - 1) >> 1);
}

// rhs_rb of one slot from rhs: one thread per interior cell of a level
template <int DIM, typename R>
AFG_DEV void split_rhs_body(const DevTree &T, const LevelArgs &L, const R *rhs, R *rhs_rb, int rb_stride, int gid)
{
    const int nc = L.nc;
    const int per_box = ipow(nc, DIM);
    const int b = gid / per_box;
    const int c = gid - b * per_box;
    const int i = c % nc + 1;
    const int j = (c / nc) % nc + 1;
    const int k = DIM == 3 ? c / (nc * nc) + 1 : 0;
    const int s = nc + 2;
    const int slot = L.slot0 + b;
    rhs_rb[(size_t)slot * rb_stride + colour_index<DIM>(i, j, k, nc)] =
        rhs[(size_t)slot * T.stride + i + s * (j + s * k)];
}

template <int DIM, int NC, typename R>
AFG_DEV void gsrb_body(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &own,
                       const LevelViewT<R> &par, int op_auto, int rb, int fly, int gid)
{
    const int nc = NC ? NC : L.nc;
    const int half = ipow(nc, DIM) / 2;
    const int b = gid / half;
    int i, j, k;
    colour_cell<DIM>(gid - b * half, nc, rb, i, j, k);
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    const size_t base = own.off(slot);
    R *box = own.phi + base;
    const int q = i + s * (j + s * k);
    const R center = ldv(box + q);
    R u[6];
    stencil<DIM>(T, L, own, par, slot, box, nc, i, j, k, q, fly, center, u);
    if (lsf_box(T, own, slot)) {
        box[q] = gsrb_value_lsf<DIM>(u, ldv(own.rhs + base + q), own.lsf + base, own.bval + base, q, s, L.dx2);
        return;
    }
    const bool lpld = lpld_box(T, op_auto, slot);
    if (DIM == 2 && cyl_op(T, op_auto))
        box[q] = gsrb_value_cyl(u, ldv(own.rhs + base + q), lpld ? own.eps + base : nullptr, q, s, L.dx2, lpld,
                                T.cyl_ioff[slot] + i);
    else
        box[q] = gsrb_value<DIM>(u, ldv(own.rhs + base + q), lpld ? own.eps + base : nullptr, q, s, L.dx2, lpld);
}

// ---------------------------------------------------------------------------
// one colour pass of GSRB over a level by tiles: a workgroup takes BOXES
// boxes.  Stage A copies their interiors to LDS, puts in their ghost layer the
// face ghosts that cells of the colour read -- computed with ghost_value
// (FLY) or read from memory -- and fetches rhs of the thread's cells of the
// colour; after a workgroup barrier, stage B updates those cells from LDS.
// Every global read is issued in stage A.  Results equal gsrb_body's: the
// ghosts a pass computes depend only on cells of the other colour and on the
// parent level, which the pass does not change.
template <int DIM, int NC, int V = 0>
struct GsrbTile {
    static constexpr int S = NC + 2;
    static constexpr int LEN = ipow(S, DIM);            // LDS values per box
    static constexpr int CELLS = ipow(NC, DIM);
    static constexpr int HALF = CELLS / 2;              // cells of one colour
    // shape V: workgroup size and boxes per workgroup.  0: 256 threads and
    // as many boxes as give one cell of the colour per thread (at least
    // one).  Boxes of 1024+ cells of a colour: 1: 512 threads, 2: 1024
    // threads.  256..1023 cells: 1: 128 threads on one box, 2: 512 threads on
    // 2 boxes.  Fewer: 1: twice the boxes, 2: 128 threads on half the boxes.
    static constexpr int B0 = HALF >= 256 ? 1 : 256 / HALF;
    static constexpr int THREADS = V == 0 ? 256
                                   : HALF >= 1024 ? (V == 1 ? 512 : 1024)
                                   : HALF >= 256 ? (V == 1 ? 128 : 512)
                                   : (V == 1 ? 256 : 128);
    static constexpr int BOXES = V == 0 ? B0
                                 : HALF >= 1024 ? 1
                                 : HALF >= 256 ? (V == 1 ? 1 : 2)
                                 : (V == 1 ? 2 * B0 : (B0 > 1 ? B0 / 2 : 1));
    static constexpr int FACE = DIM == 2 ? NC : NC * NC;
    static constexpr int GHOSTS = DIM * FACE;           // face ghosts of one colour
    static constexpr int PER = (BOXES * HALF + THREADS - 1) / THREADS;   // cells per thread
};



// Workgroup index -> tile index such that each of the 8 XCDs, which take
// workgroups round-robin, runs one contiguous range of tiles (neighbouring
// boxes share that XCD's L2).
AFG_HD inline int xcd_tile(int blk, int n_blk)
{
    const int q = n_blk >> 3, r = n_blk & 7;
    const int x = blk & 7, y = blk >> 3;
    return x * q + (x < r ? x : r) + y;
}

// `v` holds the level (its phi the pool read by this pass), `vp` the parent
// level (for the refinement-boundary ghosts).
template <int DIM, int NC, int FLY, int V, typename R>
AFG_DEV void tile_stage_a(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const LevelViewT<R> &vp,
                          R *lds, int box0, int rb, int tid, R *rhs)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
    // rhs of this thread's cells of the colour
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e < G::BOXES * G::HALF && box0 + b < L.n) {
            const int c = e - b * G::HALF;
            if (DIM == 3 && v.rhs_rb) {   // rb counts passes; odd passes take colour 1
                rhs[r] = v.rhs_rb[(size_t)(L.slot0 + box0 + b) * v.rb_stride + ((rb + 1) & 1) * G::HALF + c];
            } else {
                int i, j, k;
                colour_cell<DIM>(c, NC, rb, i, j, k);
                rhs[r] = v.rhs[v.off(L.slot0 + box0 + b) + i + S * (j + S * k)];
            }
        }
    }
    // interiors
    for (int e = tid; e < G::BOXES * G::CELLS; e += G::THREADS) {
        const int b = e / G::CELLS;
        if (box0 + b >= L.n) break;
        const int c = e - b * G::CELLS;
        const int i = c % NC + 1, j = (c / NC) % NC + 1, k = DIM == 3 ? c / (NC * NC) + 1 : 0;
        const int q = i + S * (j + S * k);
        lds[b * G::LEN + q] = v.phi[v.off(L.slot0 + box0 + b) + q];
    }
    // face ghosts next to cells of the colour
    constexpr int FH = G::FACE / 2;
    for (int e = tid; e < G::BOXES * G::GHOSTS; e += G::THREADS) {
        const int b = e / G::GHOSTS;
        if (box0 + b >= L.n) break;
        const int r = e - b * G::GHOSTS;
        const int f = r / FH, t = r - f * FH;
        const int dn = f >> 1, high = f & 1;
        const int m = high ? NC : 1;
        int t1, t2;
        if (DIM == 2) {
            t1 = 2 * t + 1 + ((rb + m + 1) & 1);
            t2 = 0;
        } else {
            t2 = t / (NC / 2) + 1;
            t1 = 2 * (t % (NC / 2)) + 1 + ((rb + m + t2 + 1) & 1);
        }
        const int i = dn == 0 ? m : t1;
        const int j = dn == 1 ? m : (dn == 0 ? t1 : t2);
        const int k = DIM == 3 ? (dn == 2 ? m : t2) : 0;
        const int gi = dn == 0 ? (high ? NC + 1 : 0) : i;
        const int gj = dn == 1 ? (high ? NC + 1 : 0) : j;
        const int gk = dn == 2 ? (high ? NC + 1 : 0) : k;
        const int gq = gi + S * (gj + S * gk);
        const int slot = L.slot0 + box0 + b;
        const size_t base = v.off(slot);
        R g;
        if (FLY)
            g = ghost_value<DIM>(T, v, vp, L.pnc, L.dr, slot, f, NC, i, j, k,
                                 T.nb[slot * (2 * DIM) + f] >= 0 ? R(0) : v.phi[base + i + S * (j + S * k)]);
        else
            g = v.phi[base + gq];
        lds[b * G::LEN + gq] = g;
    }
}

template <int DIM, int NC, int V, typename R>
AFG_DEV void tile_stage_b(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v,
                          const R *lds, int box0, int op_auto, int rb, int tid, const R *rhs)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e >= G::BOXES * G::HALF || box0 + b >= L.n) break;
        int i, j, k;
        colour_cell<DIM>(e - b * G::HALF, NC, rb, i, j, k);
        const int q = i + S * (j + S * k);
        const R *x = lds + b * G::LEN + q;
        R u[6];
        u[0] = x[-1];
        u[1] = x[1];
        u[2] = x[-S];
        u[3] = x[S];
        u[4] = DIM == 3 ? x[-S * S] : R(0);
        u[5] = DIM == 3 ? x[S * S] : R(0);
        const int slot = L.slot0 + box0 + b;
        const size_t base = v.off(slot);
        const bool lpld = lpld_box(T, op_auto, slot);
        v.phi[base + q] = gsrb_value<DIM>(u, rhs[r], lpld ? v.eps + base : nullptr, q, S, L.dx2, lpld);
    }
}

// ---------------------------------------------------------------------------
// Colour passes of a relaxation block on a colour-split copy of phi (3D
// tiles).  A pass needs the other colour's cells and writes its own colour;
// in the box arrays those are every other value of every interior row, so a
// pass reads and writes whole rows.  Within a block of n >= 2 passes on one
// level (consecutive launches that nothing else on the level comes between)
// the level's phi is also kept per slot without ghosts, colour 1 then colour
// 2 in colour_cell order (`phi_rb`, the layout of rhs_rb):
//   first pass  reads the box arrays (as k_gsrb_tile) and writes both colours
//               to phi_rb (its colour updated, the other as read);
//   middle      reads both colours of its boxes from phi_rb and writes its
//               colour there;
//   last        reads phi_rb and writes every interior cell to the box arrays.
// Face ghosts of the middle and last passes are computed as ghost_value
// computes them, reading the cells of this level from phi_rb.  The box arrays
// hold the state after the block, the ghost layer as before it, as after the
// per-pass kernels; phi_rb is only read within the block that wrote it.

// ghost_value on a level whose cells are read from the colour-split copy
// `rbp` (rb_stride values per slot); the parent level from `par` as ever.
// Expressions and their order are ghost_value's.
// nbv = T.nb of the face, p = T.parent and co = T.choff of the slot, loaded
// by the caller.
template <int DIM, typename R>
AFG_DEV R ghost_value_split(const DevTree &T, int nbv, int p, int co, const R *rbp, int rb_stride,
                            const LevelViewT<R> &par, int pnc, double dr, int slot, int f, int nc, int i, int j,
                            int k, R center)
{
    const int dn = f >> 1, high = f & 1;
    if (nbv >= 0) {
        const int m = high ? 1 : nc;
        const int a0 = dn == 0 ? m : i, a1 = dn == 1 ? m : j, a2 = dn == 2 ? m : k;
        return ldv(rbp + (size_t)nbv * rb_stride + colour_index<DIM>(a0, a1, a2, nc));
    }
    if (nbv == NB_REFINE) {
        const int gv = high ? nc + 1 : 0;
        const int g0 = dn == 0 ? gv : i, g1 = dn == 1 ? gv : j, g2 = dn == 2 ? gv : k;
        const int ps = pnc + 2;
        const int pi = (co & 1023) + ((g0 + 1) >> 1);
        const int pj = ((co >> 10) & 1023) + ((g1 + 1) >> 1);
        const int pk = DIM == 3 ? ((co >> 20) & 1023) + ((g2 + 1) >> 1) : 0;
        const R P = ldv(par.phi + par.off(p) + pi + ps * (pj + ps * pk));
        const R *own = rbp + (size_t)slot * rb_stride;
        const int inward = high ? -1 : 1;
        if (DIM == 2) {
            const int di = dn == 0 ? inward : -1 + 2 * (i & 1);
            const int dj = dn == 1 ? inward : -1 + 2 * (j & 1);
            const R a = ldv(own + colour_index<DIM>(i + di, j, k, nc));
            const R b = ldv(own + colour_index<DIM>(i, j + dj, k, nc));
            return R(0.5) * P + center - R(0.25) * (a + b);
        } else {
            const int d0 = dn == 0 ? inward : -1 + 2 * (i & 1);
            const int d1 = dn == 1 ? inward : -1 + 2 * (j & 1);
            const int d2 = dn == 2 ? inward : -1 + 2 * (k & 1);
            const R d = ldv(own + colour_index<DIM>(i + d0, j + d1, k + d2, nc));
            return R(0.5) * P + R(0.75) * center - R(0.25) * d;
        }
    }
    const int rec = NB_BC0 - nbv;
    const int voff = T.bc_voff[rec];
    R b = 0;
    if (voff >= 0) {
        const int t1 = dn == 0 ? j : i;
        const int t2 = dn == 2 ? j : k;
        b = R(T.bc_vals[voff + (t1 - 1) + (DIM == 3 ? nc * (t2 - 1) : 0)]);
    }
    R c1, c2;
    if (T.bc_type[rec] == -1) {   // a5_bc_dirichlet
        c1 = -1;
        c2 = 2;
    } else {                      // a5_bc_neumann
        c1 = 1;
        c2 = R(dr * (high ? 1.0 : -1.0));
    }
    return c2 * b + c1 * center;
}

// Stage A of the middle and last passes: rhs of the thread's cells, the other
// colour of the tile's boxes from phi_rb into LDS (the only interior cells
// the updates read), and the face ghosts next to cells of colour rb, computed
// from phi_rb (the cell next to the ghost is read only where ghost_value
// uses it: at refinement and physical boundaries).  The loads go out in the
// order of their dependences: the topology of the thread's ghosts, then rhs
// and the interior cells, then the ghosts' values (which need the topology);
// the LDS stores come last, so a thread waits for one chain of two loads.
template <int DIM, int NC, int V, typename R>
AFG_DEV void split_stage_a(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const R *rbp, R *lds,
                           int box0, int rb, int tid, R *rhs)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
    constexpr int FH = G::FACE / 2;
    constexpr int NG = (G::BOXES * G::GHOSTS + G::THREADS - 1) / G::THREADS;   // ghosts per thread
    constexpr int NI = (G::BOXES * G::HALF + G::THREADS - 1) / G::THREADS;     // interior cells per thread
    // the face of ghost e of the tile and its box, false past the level's end
    auto ghost_at = [&](int e, int &b, int &f, int &t) {
        b = e / G::GHOSTS;
        const int r = e - b * G::GHOSTS;
        f = r / FH;
        t = r - f * FH;
        return e < G::BOXES * G::GHOSTS && box0 + b < L.n;
    };
    int nbv[NG] = {}, par[NG] = {}, co[NG] = {};
#pragma unroll
    for (int g = 0; g < NG; g++) {
        int b, f, t;
        if (ghost_at(tid + g * G::THREADS, b, f, t)) {
            const int slot = L.slot0 + box0 + b;
            nbv[g] = T.nb[slot * (2 * DIM) + f];
            par[g] = T.parent[slot];
            co[g] = T.choff[slot];
        }
    }
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e < G::BOXES * G::HALF && box0 + b < L.n)
            rhs[r] = v.rhs_rb[(size_t)(L.slot0 + box0 + b) * v.rb_stride + ((rb + 1) & 1) * G::HALF + (e - b * G::HALF)];
    }
    // interior cells of the other colour: split index c holds colour 1
    // (i + j + k odd) for c < HALF, colour 2 above
    R xin[NI] = {};
#pragma unroll
    for (int r = 0; r < NI; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e < G::BOXES * G::HALF && box0 + b < L.n)
            xin[r] = ldv(rbp + (size_t)(L.slot0 + box0 + b) * v.rb_stride + (rb & 1) * G::HALF + (e - b * G::HALF));
    }
    // face ghosts next to cells of the colour (the enumeration of tile_stage_a)
    R gv[NG];
    int gpos[NG];
#pragma unroll
    for (int g = 0; g < NG; g++) {
        int b, f, t;
        gpos[g] = -1;
        if (!ghost_at(tid + g * G::THREADS, b, f, t)) continue;
        const int dn = f >> 1, high = f & 1;
        const int m = high ? NC : 1;
        const int t2 = t / (NC / 2) + 1;
        const int t1 = 2 * (t % (NC / 2)) + 1 + ((rb + m + t2 + 1) & 1);
        const int i = dn == 0 ? m : t1;
        const int j = dn == 1 ? m : (dn == 0 ? t1 : t2);
        const int k = dn == 2 ? m : t2;
        const int gi = dn == 0 ? (high ? NC + 1 : 0) : i;
        const int gj = dn == 1 ? (high ? NC + 1 : 0) : j;
        const int gk = dn == 2 ? (high ? NC + 1 : 0) : k;
        const int slot = L.slot0 + box0 + b;
        const R center = nbv[g] >= 0 ? R(0) : ldv(rbp + (size_t)slot * v.rb_stride + colour_index<DIM>(i, j, k, NC));
        gv[g] = ghost_value_split<DIM>(T, nbv[g], par[g], co[g], rbp, v.rb_stride, v, L.pnc, L.dr, slot, f, NC, i, j,
                                       k, center);
        gpos[g] = b * G::LEN + gi + S * (gj + S * gk);
    }
#pragma unroll
    for (int r = 0; r < NI; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e < G::BOXES * G::HALF && box0 + b < L.n) {
            int i, j, k;
            colour_cell<DIM>(e - b * G::HALF, NC, rb + 1, i, j, k);
            lds[b * G::LEN + i + S * (j + S * k)] = xin[r];
        }
    }
#pragma unroll
    for (int g = 0; g < NG; g++)
        if (gpos[g] >= 0) lds[gpos[g]] = gv[g];
}

// Stage B of the split passes: colour rb updated from LDS as tile_stage_b
// does; MODE 0 (first) and 1 (middle) write it to phi_rb, the first pass
// also the other colour as read; MODE 2 (last) writes both colours to the box
// arrays; MODE 3 (last, with the ghost fill) also puts colour rb into LDS for
// split_fill_push.
template <int DIM, int NC, int V, int MODE, typename R>
AFG_DEV void split_stage_b(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, R *rbp, R *lds,
                           int box0, int op_auto, int rb, int tid, const R *rhs)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e >= G::BOXES * G::HALF || box0 + b >= L.n) break;
        const int c = e - b * G::HALF;
        int i, j, k;
        colour_cell<DIM>(c, NC, rb, i, j, k);
        const int q = i + S * (j + S * k);
        const R *x = lds + b * G::LEN + q;
        R u[6];
        u[0] = x[-1];
        u[1] = x[1];
        u[2] = x[-S];
        u[3] = x[S];
        u[4] = DIM == 3 ? x[-S * S] : R(0);
        u[5] = DIM == 3 ? x[S * S] : R(0);
        const int slot = L.slot0 + box0 + b;
        const size_t base = v.off(slot);
        const bool lpld = lpld_box(T, op_auto, slot);
        const R nv = gsrb_value<DIM>(u, rhs[r], lpld ? v.eps + base : nullptr, q, S, L.dx2, lpld);
        if (MODE >= 2)
            v.phi[base + q] = nv;
        else
            rbp[(size_t)slot * v.rb_stride + ((rb + 1) & 1) * G::HALF + c] = nv;
        if (MODE == 3) lds[b * G::LEN + q] = nv;
    }
    if (MODE == 1) return;
    // the other colour, unchanged by this pass
    for (int e = tid; e < G::BOXES * G::HALF; e += G::THREADS) {
        const int b = e / G::HALF;
        if (box0 + b >= L.n) break;
        const int c = e - b * G::HALF;
        int i, j, k;
        colour_cell<DIM>(c, NC, rb + 1, i, j, k);
        const int q = i + S * (j + S * k);
        const int slot = L.slot0 + box0 + b;
        if (MODE >= 2)
            v.phi[v.off(slot) + q] = lds[b * G::LEN + q];
        else
            rbp[(size_t)slot * v.rb_stride + (rb & 1) * G::HALF + c] = lds[b * G::LEN + q];
    }
}

// The ghost fill of the level after its last pass (MODE 3), from the tile's
// final cells in LDS (both colours, after a workgroup barrier): every face
// cell goes to the ghost layer of the same-level neighbour on that face (the
// value the neighbour's fill would read there: sides_from_nb), and at
// refinement and physical boundaries the box's own ghost is formed with
// ghost_value's expressions from its LDS cells and the parent level.  Every
// face of every box of the level is then filled, as by fill_body: a face with
// a same-level neighbour by that neighbour's workgroup.  Nothing in the launch
// reads the box arrays' ghost layer (the pass reads phi_rb).
template <int DIM, int NC, int V, typename R>
AFG_DEV void split_fill_push(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const R *lds, int box0,
                             int tid)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
    constexpr int ALL = 2 * DIM * G::FACE;   // face cells of a box
    for (int e = tid; e < G::BOXES * ALL; e += G::THREADS) {
        const int b = e / ALL;
        if (box0 + b >= L.n) break;
        const int r = e - b * ALL;
        const int f = r / G::FACE, t = r - f * G::FACE;
        const int t1 = t % NC + 1, t2 = t / NC + 1;
        const int dn = f >> 1, high = f & 1;
        const int m = high ? NC : 1;
        const int i = dn == 0 ? m : t1;
        const int j = dn == 1 ? m : (dn == 0 ? t1 : t2);
        const int k = dn == 2 ? m : t2;
        const int slot = L.slot0 + box0 + b;
        const R *xb = lds + b * G::LEN;
        const R x = xb[i + S * (j + S * k)];
        const int nbv = T.nb[slot * (2 * DIM) + f];
        if (nbv >= 0) {   // the neighbour's ghost on its opposite face
            const int gm = high ? 0 : NC + 1;
            const int gi = dn == 0 ? gm : i, gj = dn == 1 ? gm : j, gk = dn == 2 ? gm : k;
            v.phi[v.off(nbv) + gi + S * (gj + S * gk)] = x;
        } else {
            const int gm = high ? NC + 1 : 0;
            const int gi = dn == 0 ? gm : i, gj = dn == 1 ? gm : j, gk = dn == 2 ? gm : k;
            v.phi[v.off(slot) + gi + S * (gj + S * gk)] =
                ghost_value<DIM>(T, v, v, L.pnc, L.dr, slot, f, NC, i, j, k, x, xb);
        }
    }
}

// ---------------------------------------------------------------------------
// Two colour passes in one launch: pass rb (colour A) and pass rb + 1
// (colour B) of a relaxation block, reading the level's phi from one pool
// and writing it to another (the two pools swap roles every pair), so that
// every global read sees the state before the pair.  The views: `v` with the
// pool read for the level, `vp` with the pool of the other levels (the
// parent's phi of the refinement-boundary ghosts).  A workgroup takes the
// boxes of a tile as k_gsrb_tile does:
//   stage A   (tile_stage_a) interiors, colour-A face ghosts and rhs;
//   phase 1   pass A on the tile's colour-A cells, in LDS; the colour-B face
//             ghosts at same-level neighbours: the neighbour's mirrored cell
//             after pass A, computed here from the pool exactly as the
//             neighbour's own workgroup computes it; those at physical
//             boundaries: the boundary formula;
//   phase 2   the colour-B ghosts at refinement boundaries (they read the
//             box's own cells after pass A, from LDS);
//   phase 3   pass B on the colour-B cells; every interior cell is written
//             to the output pool.
// The values are those of the two passes run one after the other: each
// cell and ghost is formed by the same expression from the same inputs.

// rhs of the tile's cells of colour rb, as tile_stage_a loads it
template <int DIM, int NC, int V, typename R>
AFG_DEV void pair_load_rhs(const LevelArgs &L, const LevelViewT<R> &v, int box0, int rb, int tid, R *rhs)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e < G::BOXES * G::HALF && box0 + b < L.n) {
            const int c = e - b * G::HALF;
            if (DIM == 3 && v.rhs_rb) {   // rb counts passes; odd passes take colour 1
                rhs[r] = v.rhs_rb[(size_t)(L.slot0 + box0 + b) * v.rb_stride + ((rb + 1) & 1) * G::HALF + c];
            } else {
                int i, j, k;
                colour_cell<DIM>(c, NC, rb, i, j, k);
                rhs[r] = v.rhs[v.off(L.slot0 + box0 + b) + i + S * (j + S * k)];
            }
        }
    }
}

// The value cell (i, j, k) of box y has after a colour pass on y's level
// (the pass of that cell's colour; the pool holds the state before it):
// what tile_stage_a / tile_stage_b compute for it, ghosts computed when FLY
// and read from memory otherwise.
template <int DIM, int NC, int FLY, typename R>
AFG_DEV R cell_after_pass(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const LevelViewT<R> &vp,
                          int op_auto, int y, int i, int j, int k)
{
    constexpr int S = NC + 2;
    const size_t base = v.off(y);
    const R *box = v.phi + base;
    const int q = i + S * (j + S * k);
    const int x[3] = {i, j, k};
    const int dq[6] = {-1, 1, -S, S, -S * S, S * S};
    R u[6];
#pragma unroll
    for (int m = 0; m < 2 * DIM; m++) {
        const int dn = m >> 1, high = m & 1;
        const bool inside = high ? x[dn] < NC : x[dn] > 1;
        if (inside || !FLY)
            u[m] = box[q + dq[m]];
        else
            u[m] = ghost_value<DIM>(T, v, vp, L.pnc, L.dr, y, m, NC, i, j, k, box[q]);
    }
    if (DIM == 2) u[4] = u[5] = R(0);
    const bool lpld = lpld_box(T, op_auto, y);
    const R rhs = DIM == 3 && v.rhs_rb ? v.rhs_rb[(size_t)y * v.rb_stride + colour_index<DIM>(i, j, k, NC)] : v.rhs[base + q];
    return gsrb_value<DIM>(u, rhs, lpld ? v.eps + base : nullptr, q, S, L.dx2, lpld);
}

// the colour-B face ghost e of the tile: (box b, face f, adjacent interior
// cell (i, j, k), ghost index gq), in the enumeration of tile_stage_a
template <int DIM, int NC, int V>
AFG_DEV void pair_ghost(int e, int rb, int &b, int &f, int &i, int &j, int &k, int &gq)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S, FH = G::FACE / 2;
    b = e / G::GHOSTS;
    const int r = e - b * G::GHOSTS;
    f = r / FH;
    const int t = r - f * FH;
    const int dn = f >> 1, high = f & 1;
    const int m = high ? NC : 1;
    int t1, t2;
    if (DIM == 2) {
        t1 = 2 * t + 1 + ((rb + m + 1) & 1);
        t2 = 0;
    } else {
        t2 = t / (NC / 2) + 1;
        t1 = 2 * (t % (NC / 2)) + 1 + ((rb + m + t2 + 1) & 1);
    }
    i = dn == 0 ? m : t1;
    j = dn == 1 ? m : (dn == 0 ? t1 : t2);
    k = DIM == 3 ? (dn == 2 ? m : t2) : 0;
    const int gi = dn == 0 ? (high ? NC + 1 : 0) : i;
    const int gj = dn == 1 ? (high ? NC + 1 : 0) : j;
    const int gk = dn == 2 ? (high ? NC + 1 : 0) : k;
    gq = gi + S * (gj + S * gk);
}

template <int DIM, int NC, int FLYA, int V, typename R>
AFG_DEV void pair_phase1(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const LevelViewT<R> &vp,
                         R *lds, int box0, int op_auto, int rbA, int tid, const R *rhsA)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e >= G::BOXES * G::HALF || box0 + b >= L.n) break;
        int i, j, k;
        colour_cell<DIM>(e - b * G::HALF, NC, rbA, i, j, k);
        const int q = i + S * (j + S * k);
        R *x = lds + b * G::LEN + q;
        R u[6];
        u[0] = x[-1];
        u[1] = x[1];
        u[2] = x[-S];
        u[3] = x[S];
        u[4] = DIM == 3 ? x[-S * S] : R(0);
        u[5] = DIM == 3 ? x[S * S] : R(0);
        const int slot = L.slot0 + box0 + b;
        const bool lpld = lpld_box(T, op_auto, slot);
        x[0] = gsrb_value<DIM>(u, rhsA[r], lpld ? v.eps + v.off(slot) : nullptr, q, S, L.dx2, lpld);
    }
    const int rbB = rbA + 1;
    for (int e = tid; e < G::BOXES * G::GHOSTS; e += G::THREADS) {
        int b, f, i, j, k, gq;
        pair_ghost<DIM, NC, V>(e, rbB, b, f, i, j, k, gq);
        if (box0 + b >= L.n) break;
        const int slot = L.slot0 + box0 + b;
        const int nbv = T.nb[slot * (2 * DIM) + f];
        if (nbv == NB_REFINE) continue;   // phase 2
        R g;
        if (nbv >= 0) {
            const int dn = f >> 1, high = f & 1;
            const int mm = high ? 1 : NC;   // the neighbour's mirrored cell
            g = cell_after_pass<DIM, NC, FLYA>(T, L, v, vp, op_auto, nbv, dn == 0 ? mm : i, dn == 1 ? mm : j,
                                              dn == 2 ? mm : k);
        } else {
            g = ghost_value<DIM>(T, v, vp, L.pnc, L.dr, slot, f, NC, i, j, k, lds[b * G::LEN + i + S * (j + S * k)]);
        }
        lds[b * G::LEN + gq] = g;
    }
}

template <int DIM, int NC, int V, typename R>
AFG_DEV void pair_phase2(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const LevelViewT<R> &vp,
                         R *lds, int box0, int rbA, int tid)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
    for (int e = tid; e < G::BOXES * G::GHOSTS; e += G::THREADS) {
        int b, f, i, j, k, gq;
        pair_ghost<DIM, NC, V>(e, rbA + 1, b, f, i, j, k, gq);
        if (box0 + b >= L.n) break;
        const int slot = L.slot0 + box0 + b;
        if (T.nb[slot * (2 * DIM) + f] != NB_REFINE) continue;
        R *xb = lds + b * G::LEN;
        xb[gq] = ghost_value<DIM>(T, v, vp, L.pnc, L.dr, slot, f, NC, i, j, k, xb[i + S * (j + S * k)], xb);
    }
}

template <int DIM, int NC, int V, typename R>
AFG_DEV void pair_phase3(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const R *lds, int box0,
                         int op_auto, int rbA, int tid, const R *rhsB, R *out)
{
    using G = GsrbTile<DIM, NC, V>;
    constexpr int S = G::S;
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        const int e = tid + r * G::THREADS;
        const int b = e / G::HALF;
        if (e >= G::BOXES * G::HALF || box0 + b >= L.n) break;
        const int slot = L.slot0 + box0 + b;
        const size_t base = v.off(slot);
        int i, j, k;
        colour_cell<DIM>(e - b * G::HALF, NC, rbA + 1, i, j, k);
        int q = i + S * (j + S * k);
        const R *x = lds + b * G::LEN + q;
        R u[6];
        u[0] = x[-1];
        u[1] = x[1];
        u[2] = x[-S];
        u[3] = x[S];
        u[4] = DIM == 3 ? x[-S * S] : R(0);
        u[5] = DIM == 3 ? x[S * S] : R(0);
        const bool lpld = lpld_box(T, op_auto, slot);
        out[base + q] = gsrb_value<DIM>(u, rhsB[r], lpld ? v.eps + base : nullptr, q, S, L.dx2, lpld);
        // the colour-A cell of the same index, final since phase 1
        colour_cell<DIM>(e - b * G::HALF, NC, rbA, i, j, k);
        q = i + S * (j + S * k);
        out[base + q] = lds[b * G::LEN + q];
    }
}

// residual_box (m_aX_mg.f90:545-559) over a level: tmp = rhs - A phi
template <int DIM, int NC, typename R>
AFG_DEV R residual_body(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &own,
                        const LevelViewT<R> &par, int op_auto, int gid)
{
    const int nc = NC ? NC : L.nc;
    const int per_box = ipow(nc, DIM);
    const int b = gid / per_box;
    const int c = gid - b * per_box;
    const int i = c % nc + 1;
    const int j = (c / nc) % nc + 1;
    const int k = DIM == 3 ? c / (nc * nc) + 1 : 0;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    const size_t base = own.off(slot);
    const R *box = own.phi + base;
    const int q = i + s * (j + s * k);
    const bool lpld = lpld_box(T, op_auto, slot);
    R u[6];
    stencil<DIM>(T, L, own, par, slot, box, nc, i, j, k, q, 0, ldv(box + q), u);
    const R a = lsf_box(T, own, slot)
                    ? box_op_lsf<DIM>(u, ldv(box + q), own.lsf + base, own.bval + base, q, s, L.inv_dr_sq)
                : DIM == 2 && cyl_op(T, op_auto)
                    ? box_op_cyl(u, ldv(box + q), lpld ? own.eps + base : nullptr, q, s, L.inv_dr_sq, lpld, T.cyl_ioff[slot] + i)
                    : box_op<DIM>(u, ldv(box + q), lpld ? own.eps + base : nullptr, q, s, L.inv_dr_sq, lpld);
    const R rh = DIM == 3 && own.rhs_rb ? own.rhs_rb[(size_t)slot * own.rb_stride + colour_index<DIM>(i, j, k, nc)]
                            : ldv(own.rhs + base + q);
    const R r = rh - a;
    own.tmp[base + q] = r;
    return r;
}

// residual_body, and (float) of the residual into r32 (an FP32 pool of the
// same slot layout) and r32_rb (its colour-split copy, 3D; null: none)
template <int DIM, int NC, typename R>
AFG_DEV void residual32_body(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, int op_auto, float *r32,
                             float *r32_rb, int rb_stride, int gid)
{
    const R r = residual_body<DIM, NC>(T, L, v, v, op_auto, gid);
    const int nc = NC ? NC : L.nc;
    const int per_box = ipow(nc, DIM);
    const int b = gid / per_box, c = gid - b * per_box;
    const int i = c % nc + 1, j = (c / nc) % nc + 1, k = DIM == 3 ? c / (nc * nc) + 1 : 0;
    const int s = nc + 2, slot = L.slot0 + b;
    const float f = (float)r;
    r32[v.off(slot) + i + s * (j + s * k)] = f;
    if (DIM == 3 && r32_rb) r32_rb[(size_t)slot * rb_stride + colour_index<DIM>(i, j, k, nc)] = f;
}

// update_coarse, child part (m_aX_mg.f90:444-465), fused: per coarse cell of a
// child's octant, the residual of its 2^DIM fine cells restricted into the
// parent's tmp and phi restricted into the parent's phi (a$D_restrict_box,
// m_aX_restrict.f90:67-132).  The reference stores the residual in the
// child's tmp and restores tmp afterwards; with keep_child_tmp = 0
// (set_coarse_phi_rhs, m_aX_mg.f90:495-504) the residual stays in the
// child's tmp.  With fly = 1 the residual's face neighbours are computed as
// the ghost fill after the preceding relaxation would store them.
template <int DIM, int NC, typename R, typename W = R>
AFG_DEV void uc_child_body(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &own,
                           const LevelViewT<R> &par, int op_auto, int keep_child_tmp, int fly, int gid)
{
    constexpr bool same = std::is_same<R, W>::value;
    const int nc = NC ? NC : L.nc;
    const int hn = nc >> 1;
    const int per_box = ipow(hn, DIM);
    const int b = gid / per_box;
    const int c = gid - b * per_box;
    const int ii = c % hn + 1;
    const int jj = (c / hn) % hn + 1;
    const int kk = DIM == 3 ? c / (hn * hn) + 1 : 0;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    const size_t base = own.off(slot);
    const bool lpld = lpld_box(T, op_auto, slot);
    const R *ph = own.phi + base;
    const R *e = lpld ? own.eps + base : nullptr;
    const R *rh = own.rhs + base;
    W sr = 0, sp = 0, vr[8], vp[8];
    R lc[8];
    const bool cyl = DIM == 2 && cyl_op(T, op_auto);
    const bool lsf = lsf_box(T, own, slot), lsf_rstr = lsf_rstr_box(T, own, slot);
#pragma unroll
    for (int m = 0; m < (1 << DIM); m++) {
        const int fi = 2 * ii - 1 + (m & 1);
        const int fj = 2 * jj - 1 + ((m >> 1) & 1);
        const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
        const int q = fi + s * (fj + s * fk);
        R u[6];
        stencil<DIM>(T, L, own, par, slot, ph, nc, fi, fj, fk, q, fly, ldv(ph + q), u);
        W res;
        if constexpr (same) {
            res = ldv(rh + q) - (lsf ? box_op_lsf<DIM>(u, ldv(ph + q), own.lsf + base, own.bval + base, q, s, L.inv_dr_sq)
                           : cyl ? box_op_cyl(u, ldv(ph + q), e, q, s, L.inv_dr_sq, lpld, T.cyl_ioff[slot] + fi)
                                 : box_op<DIM>(u, ldv(ph + q), e, q, s, L.inv_dr_sq, lpld));
        } else {   // the Cartesian operators only (FP32 stages)
            W uw[6];
#pragma unroll
            for (int d = 0; d < 6; d++) uw[d] = u[d];
            res = W(ldv(rh + q)) - box_op<DIM, W, R>(uw, W(ldv(ph + q)), e, q, s, L.inv_dr_sq, lpld);
        }
        if (!keep_child_tmp) own.tmp[base + q] = R(res);
        sr = m == 0 ? res : sr + res;
        sp = m == 0 ? W(ldv(ph + q)) : sp + W(ldv(ph + q));
        vr[m] = res;
        vp[m] = ldv(ph + q);
        if (lsf_rstr) lc[m] = own.lsf[base + q];
    }
    const int p = T.parent[slot];
    const int co = T.choff[slot];
    const int ps = L.pnc + 2;
    const int pq = ((co & 1023) + ii) +
                   ps * ((((co >> 10) & 1023) + jj) + ps * (DIM == 3 ? ((co >> 20) & 1023) + kk : 0));
    const size_t pb = par.off(p);
    constexpr double w = DIM == 2 ? 0.25 : 0.125;
    if constexpr (same) {
        if (lsf_rstr) {
            const R lp = par.lsf[pb + pq];
            par.tmp[pb + pq] = restrict_lsf<DIM>(vr, lc, lp);
            par.phi[pb + pq] = restrict_lsf<DIM>(vp, lc, lp);
            return;
        }
        if (DIM == 2 && T.cyl_rmin) {
            par.tmp[pb + pq] = cyl_restrict(vr, T.cyl_rmin[p], 2 * L.dr, ii);
            par.phi[pb + pq] = cyl_restrict(vp, T.cyl_rmin[p], 2 * L.dr, ii);
            return;
        }
    }
    par.tmp[pb + pq] = R(W(w) * sr);
    par.phi[pb + pq] = R(W(w) * sp);
}

// update_coarse, child part, by tiles in LDS (NC 8 and 16): the same values
// as uc_child_body.  A workgroup of THREADS threads takes BOXES boxes:
//   A  the interiors and every face ghost into LDS (a ghost computed with the
//      ghost fill's expression from the final interior when FLY, else read);
//   B  each thread computes the residual of its PER fine cells from LDS
//      (written to the child's tmp unless keep_child_tmp), and one thread per
//      coarse cell restricts phi into the parent;
//   D  the residuals replace phi in LDS;
//   E  one thread per coarse cell restricts them into the parent's tmp.
// Sums run over the 2^DIM fine cells in the order of uc_child_body.
template <int DIM, int NC>
struct UcTile {
    static constexpr int S = NC + 2;
    static constexpr int LEN = ipow(S, DIM);
    static constexpr int CELLS = ipow(NC, DIM);
    static constexpr int HN = NC / 2;
    static constexpr int COARSE = ipow(HN, DIM);
    static constexpr int THREADS = CELLS >= 4096 ? 1024 : 256;
    static constexpr int BOXES = CELLS >= THREADS ? 1 : THREADS / CELLS;
    static constexpr int PER = BOXES * CELLS / THREADS;     // fine cells per thread
    static constexpr int FACE = DIM == 2 ? NC : NC * NC;
    static constexpr int GHOSTS = 2 * DIM * FACE;
    static_assert(BOXES * CELLS % THREADS == 0 && BOXES * COARSE <= THREADS, "tile shape");
};

template <int DIM, int NC, int FLY, typename R>
AFG_DEV void uc_tile_a(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, R *lds, int box0,
                       int tid)
{
    using G = UcTile<DIM, NC>;
    constexpr int S = G::S;
    for (int e = tid; e < G::BOXES * G::CELLS; e += G::THREADS) {
        const int b = e / G::CELLS;
        if (box0 + b >= L.n) break;
        const int c = e - b * G::CELLS;
        const int i = c % NC + 1, j = (c / NC) % NC + 1, k = DIM == 3 ? c / (NC * NC) + 1 : 0;
        const int q = i + S * (j + S * k);
        lds[b * G::LEN + q] = v.phi[v.off(L.slot0 + box0 + b) + q];
    }
    for (int e = tid; e < G::BOXES * G::GHOSTS; e += G::THREADS) {
        const int b = e / G::GHOSTS;
        if (box0 + b >= L.n) break;
        const int r = e - b * G::GHOSTS;
        const int f = r / G::FACE, t = r - f * G::FACE;
        const int dn = f >> 1, high = f & 1;
        const int m = high ? NC : 1;
        const int t1 = DIM == 2 ? t + 1 : t % NC + 1;
        const int t2 = DIM == 2 ? 0 : t / NC + 1;
        const int i = dn == 0 ? m : t1;
        const int j = dn == 1 ? m : (dn == 0 ? t1 : t2);
        const int k = DIM == 3 ? (dn == 2 ? m : t2) : 0;
        const int gi = dn == 0 ? (high ? NC + 1 : 0) : i;
        const int gj = dn == 1 ? (high ? NC + 1 : 0) : j;
        const int gk = dn == 2 ? (high ? NC + 1 : 0) : k;
        const int gq = gi + S * (gj + S * gk);
        const int slot = L.slot0 + box0 + b;
        const size_t base = v.off(slot);
        R g;
        if (FLY)
            g = ghost_value<DIM>(T, v, v, L.pnc, L.dr, slot, f, NC, i, j, k,
                                 T.nb[slot * (2 * DIM) + f] >= 0 ? R(0) : v.phi[base + i + S * (j + S * k)]);
        else
            g = v.phi[base + gq];
        lds[b * G::LEN + gq] = g;
    }
}

// uc_tile_a for a level whose phi is in the colour-split copy `rbp` (3D, the
// relaxation block before it left it there): the interiors from the copy, every
// face ghost computed with ghost_value's expressions from the copy
// (ghost_value_split) and the parent level.
template <int DIM, int NC, typename R>
AFG_DEV void uc_tile_a_split(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const R *rbp, R *lds,
                             int box0, int tid)
{
    using G = UcTile<DIM, NC>;
    constexpr int S = G::S;
    constexpr int HALF = G::CELLS / 2;
    for (int e = tid; e < G::BOXES * G::CELLS; e += G::THREADS) {
        const int b = e / G::CELLS;
        if (box0 + b >= L.n) break;
        const int c = e - b * G::CELLS;
        const int hi = c >= HALF;
        int i, j, k;
        colour_cell<DIM>(c - hi * HALF, NC, hi ? 2 : 1, i, j, k);
        lds[b * G::LEN + i + S * (j + S * k)] = ldv(rbp + (size_t)(L.slot0 + box0 + b) * v.rb_stride + c);
    }
    for (int e = tid; e < G::BOXES * G::GHOSTS; e += G::THREADS) {
        const int b = e / G::GHOSTS;
        if (box0 + b >= L.n) break;
        const int r = e - b * G::GHOSTS;
        const int f = r / G::FACE, t = r - f * G::FACE;
        const int dn = f >> 1, high = f & 1;
        const int m = high ? NC : 1;
        const int t1 = DIM == 2 ? t + 1 : t % NC + 1;
        const int t2 = DIM == 2 ? 0 : t / NC + 1;
        const int i = dn == 0 ? m : t1;
        const int j = dn == 1 ? m : (dn == 0 ? t1 : t2);
        const int k = DIM == 3 ? (dn == 2 ? m : t2) : 0;
        const int gi = dn == 0 ? (high ? NC + 1 : 0) : i;
        const int gj = dn == 1 ? (high ? NC + 1 : 0) : j;
        const int gk = dn == 2 ? (high ? NC + 1 : 0) : k;
        const int slot = L.slot0 + box0 + b;
        const int nbv = T.nb[slot * (2 * DIM) + f];
        const R center = nbv >= 0 ? R(0) : ldv(rbp + (size_t)slot * v.rb_stride + colour_index<DIM>(i, j, k, NC));
        lds[b * G::LEN + gi + S * (gj + S * gk)] =
            ghost_value_split<DIM>(T, nbv, T.parent[slot], T.choff[slot], rbp, v.rb_stride, v, L.pnc, L.dr, slot, f,
                                   NC, i, j, k, center);
    }
}

// fine cell of the tile -> (box, LDS index)
template <int DIM, int NC>
AFG_DEV int uc_tile_cell(int e, int &b)
{
    using G = UcTile<DIM, NC>;
    b = e / G::CELLS;
    const int c = e - b * G::CELLS;
    const int i = c % NC + 1, j = (c / NC) % NC + 1, k = DIM == 3 ? c / (NC * NC) + 1 : 0;
    return i + G::S * (j + G::S * k);
}

// restriction of LDS values of coarse cell t of the tile into `to` of the parent
template <int DIM, int NC, typename R>
AFG_DEV void uc_tile_restrict(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const R *lds,
                              int box0, int t, R *to)
{
    using G = UcTile<DIM, NC>;
    constexpr int S = G::S, HN = G::HN;
    if (t >= G::BOXES * G::COARSE) return;
    const int b = t / G::COARSE;
    if (box0 + b >= L.n) return;
    const int c = t - b * G::COARSE;
    const int ii = c % HN + 1, jj = (c / HN) % HN + 1, kk = DIM == 3 ? c / (HN * HN) + 1 : 0;
    const R *x = lds + b * G::LEN;
    R sum = 0;
#pragma unroll
    for (int m = 0; m < (1 << DIM); m++) {
        const int fi = 2 * ii - 1 + (m & 1), fj = 2 * jj - 1 + ((m >> 1) & 1);
        const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
        const R val = x[fi + S * (fj + S * fk)];
        sum = m == 0 ? val : sum + val;
    }
    const int slot = L.slot0 + box0 + b;
    const int p = T.parent[slot];
    const int co = T.choff[slot];
    const int ps = L.pnc + 2;
    const int pq = ((co & 1023) + ii) +
                   ps * ((((co >> 10) & 1023) + jj) + ps * (DIM == 3 ? ((co >> 20) & 1023) + kk : 0));
    constexpr double w = DIM == 2 ? 0.25 : 0.125;
    to[v.off(p) + pq] = R(w) * sum;
}

template <int DIM, int NC, typename R>
AFG_DEV void uc_tile_b(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &v, const R *lds, int box0,
                       int op_auto, int keep_child_tmp, int tid, R *res)
{
    using G = UcTile<DIM, NC>;
    constexpr int S = G::S;
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        int b;
        const int q = uc_tile_cell<DIM, NC>(tid + r * G::THREADS, b);
        if (box0 + b >= L.n) break;
        const int slot = L.slot0 + box0 + b;
        const size_t base = v.off(slot);
        const R *x = lds + b * G::LEN + q;
        R u[6];
        u[0] = x[-1];
        u[1] = x[1];
        u[2] = x[-S];
        u[3] = x[S];
        u[4] = DIM == 3 ? x[-S * S] : R(0);
        u[5] = DIM == 3 ? x[S * S] : R(0);
        const bool lpld = lpld_box(T, op_auto, slot);
        const R rh = DIM == 3 && v.rhs_rb ? v.rhs_rb[(size_t)slot * v.rb_stride +
                                         colour_index<DIM>(q % S, (q / S) % S, DIM == 3 ? q / (S * S) : 0, NC)]
                              : v.rhs[base + q];
        res[r] = rh - box_op<DIM>(u, x[0], lpld ? v.eps + base : nullptr, q, S, L.inv_dr_sq, lpld);
        if (!keep_child_tmp) v.tmp[base + q] = res[r];
    }
    uc_tile_restrict<DIM, NC>(T, L, v, lds, box0, tid, v.phi);
}

template <int DIM, int NC, typename R>
AFG_DEV void uc_tile_d(const LevelArgs &L, R *lds, int box0, int tid, const R *res)
{
    using G = UcTile<DIM, NC>;
#pragma unroll
    for (int r = 0; r < G::PER; r++) {
        int b;
        const int q = uc_tile_cell<DIM, NC>(tid + r * G::THREADS, b);
        if (box0 + b >= L.n) break;
        lds[b * G::LEN + q] = res[r];
    }
}

// update_coarse, parent part (m_aX_mg.f90:467-479), fused with the ghost fill
// of the parent level that precedes it: one thread per cell of the whole
// (nc+2)^DIM array of every box of the level.  Face ghost cells get the
// ghost value (fill_gc_phi); on parents, rhs = A phi + tmp on the interior,
// rhs = rhs + tmp elsewhere (a$D_box_add_cc over the whole array), then
// tmp = phi (a$D_box_copy_cc) if copy_tmp.  A phi uses the filled ghost
// values, computed on the fly.  L is the parent level, own its view, par the
// view of the level below it.  fly = 0: the face ghosts are already filled
// (fill_body), and A phi reads them.
template <int DIM, int NC, typename R, typename W = R>
AFG_DEV void uc_parent_body(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &own,
                            const LevelViewT<R> &par, int op_auto, int copy_tmp, int gid, int fly = 1)
{
    const int nc = NC ? NC : L.nc;
    const int s = nc + 2;
    const int per_box = ipow(s, DIM);
    const int b = gid / per_box;
    const int q = gid - b * per_box;
    const int i = q % s;
    const int j = (q / s) % s;
    const int k = DIM == 3 ? q / (s * s) : 0;
    const int slot = L.slot0 + b;
    const size_t base = own.off(slot);
    R *box = own.phi + base;
    const int oi = (i < 1 || i > nc), oj = (j < 1 || j > nc), ok = DIM == 3 ? (k < 1 || k > nc) : 0;
    const int n_out = oi + oj + ok;
    const bool is_par = T.is_parent[slot] != 0;
    if (n_out == 0) {
        if (!is_par) return;
        const bool lpld = lpld_box(T, op_auto, slot);
        R u[6];
        stencil<DIM>(T, L, own, par, slot, box, nc, i, j, k, q, fly, ldv(box + q), u);
        R tau;
        if constexpr (std::is_same<R, W>::value) {
            const R a = lsf_box(T, own, slot)
                            ? box_op_lsf<DIM>(u, ldv(box + q), own.lsf + base, own.bval + base, q, s, L.inv_dr_sq)
                        : DIM == 2 && cyl_op(T, op_auto)
                            ? box_op_cyl(u, ldv(box + q), lpld ? own.eps + base : nullptr, q, s, L.inv_dr_sq, lpld,
                                         T.cyl_ioff[slot] + i)
                            : box_op<DIM>(u, ldv(box + q), lpld ? own.eps + base : nullptr, q, s, L.inv_dr_sq, lpld);
            tau = a + ldv(own.tmp + base + q);
        } else {   // the Cartesian operators only (FP32 stages)
            W uw[6];
#pragma unroll
            for (int d = 0; d < 6; d++) uw[d] = u[d];
            const W a = box_op<DIM, W, R>(uw, W(ldv(box + q)), lpld ? own.eps + base : nullptr, q, s, L.inv_dr_sq,
                                          lpld);
            tau = R(a + W(ldv(own.tmp + base + q)));
        }
        own.rhs[base + q] = tau;
        if (DIM == 3 && own.rhs_rb) own.rhs_rb[(size_t)slot * own.rb_stride + colour_index<DIM>(i, j, k, nc)] = tau;
        if (copy_tmp) own.tmp[base + q] = ldv(box + q);
        return;
    }
    R v = ldv(box + q);
    if (n_out == 1 && fly) {   // face ghost: fill it
        const int f = oi ? (i > nc) : oj ? 2 + (j > nc) : 4 + (k > nc);
        const int ci = oi ? (i > nc ? nc : 1) : i;
        const int cj = oj ? (j > nc ? nc : 1) : j;
        const int ck = DIM == 3 ? (ok ? (k > nc ? nc : 1) : k) : 0;
        const R center = T.nb[slot * (2 * DIM) + f] >= 0 ? R(0) : ldv(box + ci + s * (cj + s * ck));
        v = ghost_value<DIM>(T, own, par, L.pnc, L.dr, slot, f, nc, ci, cj, ck, center);
        box[q] = v;
    }
    if (!is_par) return;
    own.rhs[base + q] = ldv(own.rhs + base + q) + ldv(own.tmp + base + q);
    if (copy_tmp) own.tmp[base + q] = v;
}

// Quadratic prolongation stencil of a$D_prolong2 (m_aX_prolong.f90:281-343)
// on coarse values c(q) supplied by `get`; the value of child octant m is
// (((f0 -/+ fx) -/+ fy) -/+ fz) + f2.
template <int DIM, typename R>
struct Prolong2 {
    R f0, fx, fy, fz, f2;
    template <typename G>
    AFG_DEV Prolong2(G get, int pq, int ps)
    {
        f0 = get(pq);
        const R xm = get(pq - 1), xp = get(pq + 1);
        const R ym = get(pq - ps), yp = get(pq + ps);
        fx = R(0.125) * (xp - xm);
        fy = R(0.125) * (yp - ym);
        const R fxx = R(0.03125) * (xm - 2 * f0 + xp);
        const R fyy = R(0.03125) * (ym - 2 * f0 + yp);
        if (DIM == 3) {
            const R zm = get(pq - ps * ps), zp = get(pq + ps * ps);
            fz = R(0.125) * (zp - zm);
            const R fzz = R(0.03125) * (zm - 2 * f0 + zp);
            f2 = fxx + fyy + fzz;
        } else {
            fz = 0;
            f2 = fxx + fyy;
        }
    }
    AFG_DEV R at(int m) const
    {
        R v = (m & 1) ? f0 + fx : f0 - fx;
        v = ((m >> 1) & 1) ? v + fy : v - fy;
        if (DIM == 3) v = ((m >> 2) & 1) ? v + fz : v - fz;
        return v + f2;
    }
};

// correct_children, child part: mg%box_corr for each child.  Untagged boxes:
// mg$D_box_corr_lpl = a$D_prolong2(add) of the parent's correction into phi
// (m_aX_mg.f90:708-715); tagged boxes under the auto operators:
// mg$D_box_corr_lpld (m_aX_mg.f90:918-985).  The correction is the parent's
// tmp after tmp = phi - tmp; with fly_diff = 1 that difference is formed
// here from the parent's phi and tmp (the parent's tmp is then updated by a
// later stage).  One thread per coarse cell of the child's octant.
// With rbp, the level's phi is the colour-split copy (rbp, own.rb_stride per
// slot) in place of the box arrays.
template <int DIM, int NC, typename R>
AFG_DEV void correct_body(const DevTree &T, const LevelArgs &L, const LevelViewT<R> &own,
                          const LevelViewT<R> &par, int op_auto, int fly_diff, int gid, R *rbp = nullptr)
{
    const int nc = NC ? NC : L.nc;
    const int hn = nc >> 1;
    const int per_box = ipow(hn, DIM);
    const int b = gid / per_box;
    const int c = gid - b * per_box;
    const int ii = c % hn + 1;
    const int jj = (c / hn) % hn + 1;
    const int kk = DIM == 3 ? c / (hn * hn) + 1 : 0;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    const int p = T.parent[slot];
    const int co = T.choff[slot];
    const int ps = L.pnc + 2;
    const int ic = (co & 1023) + ii, jc = ((co >> 10) & 1023) + jj;
    const int kc = DIM == 3 ? ((co >> 20) & 1023) + kk : 0;
    const int pq = ic + ps * (jc + ps * kc);
    const size_t pb = par.off(p);
    const R *pt = par.tmp + pb;
    const R *pp = par.phi + pb;
    auto corr = [&](int q) -> R { return fly_diff ? ldv(pp + q) - ldv(pt + q) : ldv(pt + q); };
    R *cph = own.phi + own.off(slot);
    R *crb = rbp ? rbp + (size_t)slot * own.rb_stride : nullptr;
    // fine cell (fi, fj, fk) of the box
    auto cell = [&](int fi, int fj, int fk) -> R * {
        return crb ? crb + colour_index<DIM>(fi, fj, fk, nc) : cph + fi + s * (fj + s * fk);
    };
    if (lsf_box(T, own, slot)) {
        // mg$D_box_corr_lpllsf (m_aX_mg.f90:1019-1102): the correction from the
        // closest coarse cell and the one-but-closest in each direction, with
        // boundary value 0 where the level set changes sign
        const R *pl = par.lsf + pb;
        const R *cl = own.lsf + own.off(slot);
#pragma unroll
        for (int m = 0; m < (1 << DIM); m++) {
            const int sx = m & 1, sy = (m >> 1) & 1, sz = (m >> 2) & 1;
            const int fi = 2 * ii - 1 + sx, fj = 2 * jj - 1 + sy;
            const int fk = DIM == 3 ? 2 * kk - 1 + sz : 0;
            const int fq = fi + s * (fj + s * fk);
            const int qb[4] = {pq, pq + (sx ? 1 : -1), pq + (sy ? ps : -ps), pq + (sz ? ps * ps : -ps * ps)};
            R d[5], v[5];
#pragma unroll
            for (int t = 0; t <= DIM; t++) lsf_dist_val(cl[fq], R(0), pl[qb[t]], corr(qb[t]), R(0), d[t + 1], v[t + 1]);
            R c[5];
            if (DIM == 2) {
                c[1] = 2 * d[2] * d[3];
                c[2] = d[1] * d[3];
                c[3] = d[1] * d[2];
            } else {
                c[1] = d[2] * d[3] * d[4];
                c[2] = d[1] * d[3] * d[4];
                c[3] = d[1] * d[2] * d[4];
                c[4] = d[1] * d[2] * d[3];
            }
            R sv = 0, sc = 0;   // sum(c * val) / sum(c), each sum from zero
#pragma unroll
            for (int t = 1; t <= DIM + 1; t++) {
                sv += c[t] * v[t];
                sc += c[t];
            }
            cph[fq] = ldv(cph + fq) + sv / sc;
        }
    } else if (lpld_box(T, op_auto, slot)) {
        const R *pe = par.eps + pb;
        constexpr double third = 1 / 3.0;
        const R u0 = corr(pq), a0 = pe[pq];
#pragma unroll
        for (int m = 0; m < (1 << DIM); m++) {
            const int sx = m & 1, sy = (m >> 1) & 1, sz = (m >> 2) & 1;
            const int fi = 2 * ii - 1 + sx, fj = 2 * jj - 1 + sy;
            const int fk = DIM == 3 ? 2 * kk - 1 + sz : 0;
            // one-but-closest coarse cell: fine index odd -> -1, even -> +1
            // (m_aX_mg.f90:944-958)
            const int qn[3] = {pq + (sx ? 1 : -1), pq + (sy ? ps : -ps),
                               pq + (sz ? ps * ps : -ps * ps)};
            R *t = cell(fi, fj, fk);
            if (DIM == 2) {
                R sum = (a0 * u0 + pe[qn[0]] * corr(qn[0])) / (a0 + pe[qn[0]]);
                sum += (a0 * u0 + pe[qn[1]] * corr(qn[1])) / (a0 + pe[qn[1]]);
                *t = ldv(t) + R(0.5) * sum;
            } else {
                R sum = 0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const R v = (a0 * u0 + pe[qn[d]] * (R(1.5) * corr(qn[d]) - R(0.5) * u0)) /
                                (a0 + pe[qn[d]]);
                    sum = d == 0 ? v : sum + v;
                }
                *t = ldv(t) + R(third) * sum;
            }
        }
    } else {
        const Prolong2<DIM, R> P2(corr, pq, ps);
#pragma unroll
        for (int m = 0; m < (1 << DIM); m++) {
            const int fi = 2 * ii - 1 + (m & 1), fj = 2 * jj - 1 + ((m >> 1) & 1);
            const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
            R *t = cell(fi, fj, fk);
            *t = P2.at(m) + ldv(t);
        }
    }
}

// a$D_restrict_box (m_aX_restrict.f90:67-132) of every box of a level into
// its parent: one thread per coarse cell.
// With lsf set, children selected by rstr_mode (mg%box_rstr, as in
// lsf_rstr_box) take mg$D_box_rstr_lpllsf.
template <int DIM, int NC, typename R>
AFG_DEV void restrict_body(const DevTree &T, const LevelArgs &L, const R *from, R *to, int gid,
                           const R *lsf = nullptr, int rstr_mode = 0)
{
    const int nc = NC ? NC : L.nc;
    const int hn = nc >> 1;
    const int per_box = ipow(hn, DIM);
    const int b = gid / per_box;
    const int c = gid - b * per_box;
    const int ii = c % hn + 1;
    const int jj = (c / hn) % hn + 1;
    const int kk = DIM == 3 ? c / (hn * hn) + 1 : 0;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    const R *cb = from + (size_t)slot * T.stride;
    const bool lsf_rstr = lsf && (rstr_mode == 2 || (rstr_mode == 1 && T.tag[slot] == TAG_LSF));
    R sum = 0, v8[8], lc[8];
#pragma unroll
    for (int m = 0; m < (1 << DIM); m++) {
        const int fi = 2 * ii - 1 + (m & 1), fj = 2 * jj - 1 + ((m >> 1) & 1);
        const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
        const R v = cb[fi + s * (fj + s * fk)];
        sum = m == 0 ? v : sum + v;
        v8[m] = v;
        if (lsf_rstr) lc[m] = lsf[(size_t)slot * T.stride + fi + s * (fj + s * fk)];
    }
    const int p = T.parent[slot];
    const int co = T.choff[slot];
    const int ps = L.pnc + 2;
    const int pq = ((co & 1023) + ii) +
                   ps * ((((co >> 10) & 1023) + jj) + ps * (DIM == 3 ? ((co >> 20) & 1023) + kk : 0));
    constexpr double w = DIM == 2 ? 0.25 : 0.125;
    if (lsf_rstr)
        to[(size_t)p * T.stride + pq] = restrict_lsf<DIM>(v8, lc, lsf[(size_t)p * T.stride + pq]);
    else if (DIM == 2 && T.cyl_rmin)
        to[(size_t)p * T.stride + pq] = cyl_restrict(v8, T.cyl_rmin[p], 2 * L.dr, ii);
    else
        to[(size_t)p * T.stride + pq] = R(w) * sum;
}

// a$D_prolong2 (m_aX_prolong.f90:243-348) from the parent into every box of
// a level; add = 0 overwrites the interior (the reference zeroes it first and
// adds, so the stored value is expr + 0).
template <int DIM, int NC, typename R>
AFG_DEV void prolong2_body(const DevTree &T, const LevelArgs &L, const R *from, R *to, int add, int gid)
{
    const int nc = NC ? NC : L.nc;
    const int hn = nc >> 1;
    const int per_box = ipow(hn, DIM);
    const int b = gid / per_box;
    const int c = gid - b * per_box;
    const int ii = c % hn + 1;
    const int jj = (c / hn) % hn + 1;
    const int kk = DIM == 3 ? c / (hn * hn) + 1 : 0;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    const int p = T.parent[slot];
    const int co = T.choff[slot];
    const int ps = L.pnc + 2;
    const int pq = ((co & 1023) + ii) +
                   ps * ((((co >> 10) & 1023) + jj) + ps * (DIM == 3 ? ((co >> 20) & 1023) + kk : 0));
    const R *pb = from + (size_t)p * T.stride;
    const Prolong2<DIM, R> P2([&](int q) -> R { return pb[q]; }, pq, ps);
    R *cb = to + (size_t)slot * T.stride;
#pragma unroll
    for (int m = 0; m < (1 << DIM); m++) {
        const int fi = 2 * ii - 1 + (m & 1), fj = 2 * jj - 1 + ((m >> 1) & 1);
        const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
        R *t = &cb[fi + s * (fj + s * fk)];
        const R old = add ? *t : R(0);
        *t = P2.at(m) + old;
    }
}

// ---------------------------------------------------------------------------
// tree update (the data side of a$D_adjust_refinement, m_aX_core.f90:452-541)

// Whole box arrays of a new pool: slot b takes the array of old slot src[b]
// (the same box before the update), or zeros for a box the update added.
// Two values per thread: a box array holds (nc+2)^dim values with nc even,
// so every array starts on a 16-byte boundary.
struct alignas(16) Pair64 {
    double x, y;
};

AFG_DEV void gather_body(int stride, const int32_t *src, const double *from, double *to, long long gid)
{
    const int half = stride >> 1;
    const int b = (int)(gid / half);
    const int q = (int)(gid - (long long)b * half);
    const int o = src[b];
    Pair64 v = {0.0, 0.0};
    if (o >= 0) v = reinterpret_cast<const Pair64 *>(from + (size_t)o * stride)[q];
    reinterpret_cast<Pair64 *>(to + (size_t)b * stride)[q] = v;
}

// a$D_prolong0 / a$D_prolong1 / a$D_prolong2 (m_aX_prolong.f90:39-108,
// 130-219, 243-348) from the parent of box `slot` (values `from`) into the
// box's interior (values `to`), for coarse cell c of the box's octant.
// add = 0: orders 1 and 2 zero the interior and add, so the stored value is
// expr + 0, and order 0 stores the parent's value; add = 1: old + value
// (order 0: old + parent).  Order 1 sums f0 + f_x + f_y (+ f_z) + old left to
// right, f0 = 0.5 (2D) or 0.25 (3D) times the parent's cell and f_x.. = 0.25
// times the neighbour towards the child cell.
template <int DIM, int ORDER, typename R>
AFG_DEV void prolong_cell(const DevTree &T, int nc, int pnc, int slot, const R *from, R *to, int add, int c)
{
    const int hn = nc >> 1;
    const int ii = c % hn + 1;
    const int jj = (c / hn) % hn + 1;
    const int kk = DIM == 3 ? c / (hn * hn) + 1 : 0;
    const int s = nc + 2;
    const int co = T.choff[slot];
    const int ps = pnc + 2;
    const int pq = ((co & 1023) + ii) +
                   ps * ((((co >> 10) & 1023) + jj) + ps * (DIM == 3 ? ((co >> 20) & 1023) + kk : 0));
    const R *pb = from + (size_t)T.parent[slot] * T.stride;
    R *cb = to + (size_t)slot * T.stride;
    auto at = [&](int m) -> R & {
        const int fi = 2 * ii - 1 + (m & 1), fj = 2 * jj - 1 + ((m >> 1) & 1);
        const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
        return cb[fi + s * (fj + s * fk)];
    };
    if constexpr (ORDER == 0) {
        const R v = pb[pq];
#pragma unroll
        for (int m = 0; m < (1 << DIM); m++) at(m) = add ? at(m) + v : v;
    } else if constexpr (ORDER == 1) {
        const int step[3] = {1, ps, ps * ps};
        const R f0 = R(DIM == 2 ? 0.5 : 0.25) * pb[pq];
        R lo[3], hi[3];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            lo[d] = R(0.25) * pb[pq - step[d]];
            hi[d] = R(0.25) * pb[pq + step[d]];
        }
#pragma unroll
        for (int m = 0; m < (1 << DIM); m++) {
            R v = f0;
#pragma unroll
            for (int d = 0; d < DIM; d++) v = v + (((m >> d) & 1) ? hi[d] : lo[d]);
            at(m) = v + (add ? at(m) : R(0));
        }
    } else {
        const Prolong2<DIM, R> P2([&](int q) -> R { return pb[q]; }, pq, ps);
#pragma unroll
        for (int m = 0; m < (1 << DIM); m++) at(m) = P2.at(m) + (add ? at(m) : R(0));
    }
}

// The boxes a tree update added, listed in `slots` (all with nc cells, their
// parents with pnc): variable `var` prolonged from the parents, add = 0.
template <int DIM, int ORDER, typename R>
AFG_DEV void prolong_list_body(const DevTree &T, int nc, int pnc, const int32_t *slots, R *var, int gid)
{
    const int per_box = ipow(nc >> 1, DIM);
    const int b = gid / per_box;
    prolong_cell<DIM, ORDER>(T, nc, pnc, slots[b], (const R *)var, var, 0, gid - b * per_box);
}

// ---------------------------------------------------------------------------
// ghost cells of any variable (a$D_gc_tree, m_aX_gc.f90:28-69) and whole-pool
// arithmetic (a$D_box_add_cc .. a$D_box_lincomb_cc, m_aX_utils.f90:273-329)

// One face ghost cell of a level's box, as fill_body, with the
// refinement-boundary method rb and the physical boundary condition bc (0:
// the boundary table; -1 / -2: a$D_bc_dirichlet_zero / a$D_bc_neumann_zero,
// m_aX_gc.f90:361-379, i.e. that condition with b = 0 on every face).
//   gc_interp: ghost = 0.5 c1 + c2/6 + f/3 (2D), c1/3 + c2/6 + c3/6 + f/3
//   (3D), summed left to right, c1..c3 cells of the parent's neighbour next
//   to the face (c1 the one over the ghost, c2 / c3 the tangential
//   neighbours towards the ghost's sibling: -1 for odd, +1 for even cell
//   indices), f the box's cell next to the ghost; the limited form caps the
//   ghost at 2 c1.  Face z of 3D takes c2 along y and c3 along x, the other
//   faces c2 along their first tangential axis.
template <int DIM>
AFG_DEV void gc_body(const DevTree &T, const LevelArgs &L, double *var, int rb, int bc, int gid)
{
    constexpr double sixth = 1 / 6.0, third = 1 / 3.0;
    const int nc = L.nc;
    const int per_face = DIM == 2 ? nc : nc * nc;
    const int per_box = 2 * DIM * per_face;
    const int b = gid / per_box;
    const int r = gid - b * per_box;
    const int f = r / per_face;
    const int t = r - f * per_face;
    const int t1 = t % nc + 1, t2 = DIM == 3 ? t / nc + 1 : 0;
    const int dn = f >> 1, high = f & 1;
    const int in = high ? nc : 1;
    const int i = dn == 0 ? in : t1;
    const int j = dn == 1 ? in : (dn == 0 ? t1 : (DIM == 3 ? t2 : 0));
    const int k = DIM == 3 ? (dn == 2 ? in : t2) : 0;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    double *box = var + (size_t)slot * T.stride;
    const double center = box[i + s * (j + s * k)];
    const int nbv = T.nb[slot * (2 * DIM) + f];
    const int gi = dn == 0 ? (high ? nc + 1 : 0) : i;
    const int gj = dn == 1 ? (high ? nc + 1 : 0) : j;
    const int gk = dn == 2 ? (high ? nc + 1 : 0) : k;
    double g;
    if (nbv >= 0 || (nbv == NB_REFINE && rb == RB_MG) || (nbv <= NB_BC0 && bc == 0)) {
        const LevelViewT<double> v{var, nullptr, nullptr, nullptr, 0, T.stride};
        g = ghost_value<DIM>(T, v, v, L.pnc, L.dr, slot, f, nc, i, j, k, center);
    } else if (nbv <= NB_BC0) {
        double c1, c2;
        if (bc == -1) {
            c1 = -1;
            c2 = 2;
        } else {
            c1 = 1;
            c2 = L.dr * (high ? 1.0 : -1.0);
        }
        g = c2 * 0.0 + c1 * center;
    } else {
        const int co = T.choff[slot];
        const int off[3] = {co & 1023, (co >> 10) & 1023, (co >> 20) & 1023};
        const int ps = L.pnc + 2;
        const int p = T.parent[slot];
        if (rb == RB_PROLONG0) {
            const int pi = off[0] + ((gi + 1) >> 1);
            const int pj = off[1] + ((gj + 1) >> 1);
            const int pk = DIM == 3 ? off[2] + ((gk + 1) >> 1) : 0;
            g = var[(size_t)p * T.stride + pi + ps * (pj + ps * pk)];
        } else {
            const double *q = var + (size_t)T.nb[p * (2 * DIM) + f] * T.stride;
            const int ixc = high ? 1 : L.pnc;
            // tangential axes: a (first), bb (second); their cell indices
            const int ta = dn == 0 ? 1 : 0, tb = dn == 2 ? 1 : 2;
            const int ca = dn == 0 ? j : i, cb = dn == 2 ? j : k;
            const int a1 = off[ta] + ((ca + 1) >> 1), a2 = a1 + 1 - 2 * (ca & 1);
            auto cell = [&](int va, int vb) {
                int x[3];
                x[dn] = ixc;
                x[ta] = va;
                x[tb] = vb;
                return q[x[0] + ps * (x[1] + ps * (DIM == 3 ? x[2] : 0))];
            };
            double c1;
            if (DIM == 2) {
                c1 = cell(a1, 0);
                const double c2 = cell(a2, 0);
                g = 0.5 * c1 + sixth * c2 + third * center;
            } else {
                const int b1 = off[tb] + ((cb + 1) >> 1), b2 = b1 + 1 - 2 * (cb & 1);
                c1 = cell(a1, b1);
                const double ca2 = cell(a2, b1), cb2 = cell(a1, b2);
                // c2 along the first tangential axis and c3 along the second,
                // except on z faces (c2 along y, c3 along x)
                const double c2 = dn == 2 ? cb2 : ca2, c3 = dn == 2 ? ca2 : cb2;
                g = third * c1 + sixth * c2 + sixth * c3 + third * center;
            }
            if (rb == RB_INTERP_LIM && g > 2 * c1) g = 2 * c1;
        }
    }
    box[gi + s * (gj + s * gk)] = g;
}

// The second ghost layer of a box: a$D_gc2_box (m_aX_gc.f90:420-456) with
// a$D_gc2_prolong1 at refinement boundaries (:498-597) and
// a$D_bc2_neumann_zero at physical boundaries (:600-634), one thread per
// value of gc_data(nc[, nc], 2*dim) (first tangential index fastest, then the
// face), written to out + slot * gstride.  A same-level neighbour gives its
// cell nc-1 (low face) or 2 (high face); a physical boundary the box's own
// cell 2 (low) or nc-1 (high); a refinement boundary the linear
// interpolation from the parent's neighbour across the face (ghost cells of
// that box included), with the child offset of a$D_get_child_offset(box, nb).
template <int DIM>
AFG_DEV void gc2_body(const DevTree &T, const LevelArgs &L, const double *var, double *out, int gstride, int gid)
{
    const int nc = L.nc;
    const int per_face = DIM == 2 ? nc : nc * nc;
    const int per_box = 2 * DIM * per_face;
    const int b = gid / per_box;
    const int r = gid - b * per_box;
    const int f = r / per_face;
    const int t = r - f * per_face;
    const int dn = f >> 1, high = f & 1;
    const int ta = dn == 0 ? 1 : 0, tb = dn == 2 ? 1 : 2;   // tangential axes, in dimension order
    const int slot = L.slot0 + b;
    const int nbv = T.nb[slot * (2 * DIM) + f];
    int x[3] = {0, 0, 0};
    x[ta] = t % nc + 1;
    x[tb] = DIM == 3 ? t / nc + 1 : 0;
    double g;
    if (nbv >= 0 || nbv <= NB_BC0) {
        const bool own = nbv <= NB_BC0;
        x[dn] = (high != (int)own) ? 2 : nc - 1;
        const int s = nc + 2;
        const double *box = var + (size_t)(own ? slot : nbv) * T.stride;
        g = box[x[0] + s * (x[1] + s * x[2])];
    } else {
        const int p = T.parent[slot];
        const double *q = var + (size_t)T.nb[p * (2 * DIM) + f] * T.stride;
        const int co = T.choff[slot];
        int off[3] = {co & 1023, (co >> 10) & 1023, DIM == 3 ? (co >> 20) & 1023 : 0};
        off[dn] += high ? -L.pnc : L.pnc;
        x[dn] = high ? nc + 2 : -1;                         // the second ghost cell's index
        int c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) {
            c1[d] = off[d] + ((x[d] + 1) >> 1);
            c2[d] = c1[d] + 1 - 2 * (x[d] & 1);
        }
        const int ps = L.pnc + 2;
        auto at = [&](int d2) {   // the parent's neighbour at c1, dimension d2 (if >= 0) at c2
            const int y0 = d2 == 0 ? c2[0] : c1[0], y1 = d2 == 1 ? c2[1] : c1[1];
            const int y2 = DIM == 3 ? (d2 == 2 ? c2[2] : c1[2]) : 0;
            return q[y0 + ps * (y1 + ps * y2)];
        };
        if (DIM == 2)
            g = 0.5 * at(-1) + 0.25 * at(0) + 0.25 * at(1);
        else
            g = 0.25 * at(-1) + 0.25 * at(0) + 0.25 * at(1) + 0.25 * at(2);
    }
    out[(size_t)slot * gstride + r] = g;
}

// flux_from_children (m_aX_core.f90:838-943) for one face cell of a leaf
// neighbour: the mean of the 2^(dim-1) fine fluxes of the parent's child on
// that face, 0.5 (a + b) in 2D and 0.25 (((a + b) + c) + d) in 3D (first
// tangential index fastest).  On a cylindrical tree (cyl_rmin: r_min(1) per
// slot, dr the level's cell size) the y faces take the radius-weighted mean
// 0.5 (w1 a + w2 b) of m_aX_core.f90:877-887, w = 1 -/+ 0.25 dr / r at the
// neighbour's cell centre r = r_min(1) + (A - 0.5) dr (a2_cyl_child_weights,
// m_aX_t.f90:518-528).  Face arrays of a slot: fx, fy (, fz) at
// fc + slot * fstride, each (nc+1) nc^(dim-1) values with extent nc+1 along
// its own direction, Fortran order.
template <int DIM>
AFG_DEV void flux_body(const FluxRec *recs, int nc, int fstride, double *fc, int gid,
                       const double *cyl_rmin = nullptr, double dr = 0)
{
    const int per = DIM == 2 ? nc : nc * nc;
    const int r = gid / per, t = gid - r * per;
    const FluxRec R = recs[r];
    const int d = R.f >> 1, high = R.f & 1;
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const int nch = nc >> 1;
    const int A = t % nc + 1, B = DIM == 3 ? t / nc + 1 : 1;
    const int ba = (A - 1) / nch, bb = DIM == 3 ? (B - 1) / nch : 0;
    const int m = A - ba * nch, m2 = B - bb * nch;
    const int cs = R.child[ba | (bb << 1)];
    const int i = high ? nc + 1 : 1, i_nb = high ? 1 : nc + 1;
    const int len = (nc + 1) * (DIM == 3 ? nc * nc : nc);
    const int e0 = d == 0 ? nc + 1 : nc, e1 = d == 1 ? nc + 1 : nc;
    auto at = [&](int slot, int xn, int xa, int xb) -> double & {
        int x[3];
        x[d] = xn;
        x[ta] = xa;
        x[tb] = xb;
        return fc[(size_t)slot * fstride + d * len + (x[0] - 1) + e0 * ((x[1] - 1) + e1 * (DIM == 3 ? x[2] - 1 : 0))];
    };
    double v;
    if (DIM == 2) {
        if (cyl_rmin && d == 1) {
            const double r = cyl_rmin[R.nb] + (A - 0.5) * dr;
            const double tw = 0.25 * dr / r;
            const double w1 = 1 - tw, w2 = 1 + tw;
            v = 0.5 * (w1 * at(cs, i, 2 * m - 1, 1) + w2 * at(cs, i, 2 * m, 1));
        } else {
            v = 0.5 * (at(cs, i, 2 * m - 1, 1) + at(cs, i, 2 * m, 1));
        }
    } else {
        v = 0.25 * (at(cs, i, 2 * m - 1, 2 * m2 - 1) + at(cs, i, 2 * m, 2 * m2 - 1) +
                    at(cs, i, 2 * m - 1, 2 * m2) + at(cs, i, 2 * m, 2 * m2));
    }
    at(R.nb, i_nb, A, B) = v;
}

// y = y + x, y = y - x, y = a y, y = a x + b y over whole pools (ARITH_*)
AFG_DEV void arith_body(int op, double a, double b, const double *x, double *y, long long i)
{
    if (op == ARITH_ADD) y[i] = y[i] + x[i];
    else if (op == ARITH_SUB) y[i] = y[i] - x[i];
    else if (op == ARITH_TIMES) y[i] = a * y[i];
    else y[i] = a * x[i] + b * y[i];
}

// ---------------------------------------------------------------------------
// bottom of the V-cycle in one workgroup, on LDS copies of the small levels
//
// Loads phi, rhs, tmp (and eps) of levels 0..hl (relative to the lowest)
// into LDS, runs the stage sequence of mg$D_fas_vcycle below and at hl
// (gsrb_boxes, update_coarse, the base relaxation, correct_children,
// fill_gc_phi; m_aX_mg.f90:218-241) with the same thread bodies separated by
// workgroup barriers, and writes phi, rhs and tmp back.

template <int N>
struct IC {
    static constexpr int value = N;
};

template <typename F>
AFG_DEV void nc_dispatch(int nc, F &&f)
{
    switch (nc) {
    case 2: f(IC<2>()); break;
    case 4: f(IC<4>()); break;
    case 8: f(IC<8>()); break;
    case 16: f(IC<16>()); break;
    default: f(IC<0>()); break;
    }
}

// Stages of the bottom part, in order (mg$D_fas_vcycle, m_aX_mg.f90:218-241):
// per level li from hl down to 1: a relaxation block of n_down sweeps, then
// update_coarse (child part, then parent part with the parent's ghost fill);
// at level 0: a block of n_base sweeps; per level li from 1 up to hl:
// correct_children (correction, then tmp = phi - tmp on the parents), the
// ghost fill and a block of n_up sweeps.  Two forms of a block of n sweeps:
// - stored ghosts (FLY = 0, 3D): the first colour pass, then (ghost fill,
//   colour pass) for passes 2..2n, then the ghost fill: 4n steps.  Filling
//   the ghosts before a pass gives the face neighbours that the reference's
//   fill after the previous pass stores, so every pass reads them from
//   memory; the ghost layer after the block is the reference's, the last
//   fill covering every face.
// - computed ghosts (FLY = 1, 2D): the 2n colour passes, the first of a
//   down-sweep block reading the stored ghosts and every other pass computing
//   the face ghosts it reads (the values the reference's fill after the
//   previous pass stores), then the ghost fill: 2n + 1 steps.  The down-sweep
//   block has no fill (update_coarse's residual computes the face ghosts it
//   would store) and the up-sweep block none before its first pass.
// At level 0 the block is one step for all 2n passes (from the cell list,
// computing the face neighbours) and the ghost fill.
enum BottomOp { BOP_GSRB, BOP_FILL, BOP_UC_CHILD, BOP_UC_PARENT, BOP_CORRECT, BOP_DIFF, BOP_BASE, BOP_BLOCK };

// op on level li; a colour pass has its colour rb and `fly` (whether it
// computes its face ghosts); BOP_BLOCK stands for the m steps of a relaxation
// block of rb sweeps whose first pass has fly (relax_step), after a ghost fill
// of the level when `lead` is set
struct BottomStep {
    int op, li, rb, fly, m, lead;
};

// The form is fixed per dimension at compile time (measured: computing the
// face ghosts in the passes is faster in 2D; in 3D a pass with computed
// ghosts costs as much as a fill and a pass, and the fill form keeps the
// ghost code out of the passes' instruction stream).
// number of steps of a relaxation block of n sweeps
template <int FLY>
AFG_HD inline int relax_steps(int n)
{
    return FLY ? 2 * n + 1 : 4 * n;
}

// step r of a relaxation block of n sweeps; `first_fly` tells whether its
// first pass computes its face ghosts
template <int FLY>
AFG_HD inline BottomStep relax_step(int li, int r, int n, int first_fly)
{
    if (FLY) {
        if (r == 2 * n) return {BOP_FILL, li, 0, 0, 0, 0};
        return {BOP_GSRB, li, r + 1, r > 0 ? 1 : first_fly, 0, 0};
    }
    if (r == 0) return {BOP_GSRB, li, 1, 0, 0, 0};
    if (r == 4 * n - 1 || (r & 1)) return {BOP_FILL, li, 0, 0, 0, 0};
    return {BOP_GSRB, li, r / 2 + 1, 0, 0, 0};
}

// The sequence above with each relaxation block as one step (BOP_BLOCK, which
// also holds the up-sweep's fill before its first pass when FLY = 0) and the
// base relaxation with its ghost fill as one step, walked by a position
// (part, li, r) that advances without divisions: part 0 the down-sweep of
// level li from hl to 1 (r: block, child part, parent part), 1 the base level
// (r: base relaxation and fill), 2 the up-sweep of level li from 1 to hl (r:
// correction, diff, block), 3 the end.  Blocks of no steps are passed over.
struct BottomSeq {
    int part, li, r;
};

template <int FLY>
AFG_HD inline int down_block_steps(const BottomArgs &A)
{
    return FLY ? 2 * A.n_down : 4 * A.n_down;   // the down block has no fill when FLY
}

template <int FLY>
AFG_HD inline void seq_settle(const BottomArgs &A, BottomSeq &q)
{
    constexpr int NU = 3;
    for (;;) {
        if (q.part == 0) {
            if (q.li < 1) {
                q.part = 1;
                q.r = 0;
            } else if (q.r >= 3) {
                q.li--;
                q.r = 0;
            } else if (q.r == 0 && down_block_steps<FLY>(A) == 0) {
                q.r = 1;
            } else {
                return;
            }
        } else if (q.part == 1) {
            if (A.n_base > 0 && q.r < 1) return;
            q.part = 2;
            q.li = 1;
            q.r = 0;
        } else if (q.part == 2) {
            if (q.li > A.hl) {
                q.part = 3;
                return;
            }
            if (q.r >= NU) {
                q.li++;
                q.r = 0;
            } else if (q.r == NU - 1 && FLY && relax_steps<FLY>(A.n_up) == 0) {
                q.r = NU;
            } else {
                return;
            }
        } else {
            return;
        }
    }
}

template <int FLY>
AFG_HD inline BottomSeq seq_first(const BottomArgs &A)
{
    BottomSeq q{0, A.hl, 0};
    seq_settle<FLY>(A, q);
    return q;
}

template <int FLY>
AFG_HD inline void seq_next(const BottomArgs &A, BottomSeq &q)
{
    q.r++;
    seq_settle<FLY>(A, q);
}

template <int FLY>
AFG_HD inline BottomStep seq_step(const BottomArgs &A, const BottomSeq &q)
{
    if (q.part == 0) {
        if (q.r == 0) return {BOP_BLOCK, q.li, A.n_down, 0, down_block_steps<FLY>(A), 0};
        return {q.r == 1 ? BOP_UC_CHILD : BOP_UC_PARENT, q.li, 0, 0, 0, 0};
    }
    if (q.part == 1) return {BOP_BASE, 0, 0, 0, 0, 0};
    if (q.part == 2) {
        if (q.r == 0) return {BOP_CORRECT, q.li, 0, 0, 0, 0};
        if (q.r == 1) return {BOP_DIFF, q.li, 0, 0, 0, 0};
        return {BOP_BLOCK, q.li, A.n_up, FLY ? 1 : 0, relax_steps<FLY>(A.n_up), FLY ? 0 : 1};
    }
    return {0, A.hl + 1, 0, 0, 0, 0};
}

// cell counts with their own code in the bottom kernel (others: generic)
template <typename F>
AFG_DEV void bottom_nc_dispatch(int nc, F &&f)
{
    switch (nc) {
    case 2: f(IC<2>()); break;
    case 4: f(IC<4>()); break;
    case 8: f(IC<8>()); break;
    default: f(IC<0>()); break;
    }
}

// The base relaxation when every box of the lowest level has 2^DIM cells and
// physical boundaries on all faces: the 2 n_base colour passes of
// bottom_run's BOP_BASE with the same expressions, one cell per lane of a
// wavefront (64 / 2^DIM boxes at a time), the cell values in registers and
// the neighbour cell of the same box read from its lane (xshfl(v, mask): the
// value of v in lane ^ mask).  The first pass reads the stored ghosts, the
// others K + c1 phi at the boundary.  The eps weights of mg$D_box_gsrb_lpld
// are formed once, with its expressions.  Cell c of a box has coordinates
// i = (c & 1) + 1, j = ((c >> 1) & 1) + 1, k = (c >> 2) + 1; its
// descriptors are in the cell list of its colour at the rank it has among
// the box's cells of that colour.  Boxes are independent of each other
// (physical boundaries on all faces), so several wavefronts can share them:
// this one takes the groups of 64 / 2^DIM boxes wave_first, wave_first +
// wave_step, ...
// W is the arithmetic type: R, or double for an FP32 level relaxed in FP64.
template <int DIM, int AUTO, typename R, typename W, typename Xshfl>
AFG_DEV void base_wave_run(const DevTree &T, const LevelViewT<R> &own, const BottomLevels &B, const BottomArgs &A,
                           double dx2, int n_box, int lane, Xshfl xshfl, int wave_first = 0, int wave_step = 1)
{
    constexpr int NCL = 1 << DIM, H = NCL / 2, S = 4, PER_WAVE = 64 / NCL;
    const int dq[6] = {-1, 1, -S, S, -S * S, S * S};
    auto odd_of = [](int c) { return ((c & 1) + ((c >> 1) & 1) + ((c >> 2) & 1) + DIM) & 1; };
    const int c = lane % NCL;
    const int odd = odd_of(c);
    int rank = 0;
    for (int c2 = 0; c2 < c; c2++) rank += odd_of(c2) == odd;
    for (int g0 = wave_first * PER_WAVE; g0 < n_box; g0 += wave_step * PER_WAVE) {
        const int b = g0 + lane / NCL;
        const bool on = b < n_box;
        W ph = 0, rh = 0, K[3] = {0, 0, 0}, cm[6] = {0, 0, 0, 0, 0, 0}, sc = 1;
        int pos = 0, neg = 0;
        bool lpld = false;
        if (on) {
            const BaseCell &d = B.base[(odd ? 0 : B.base_n[0]) + b * H + rank];
            pos = d.pos;
            ph = W(ldv(own.phi + pos));
            rh = W(ldv(own.rhs + pos));
#pragma unroll
            for (int dn = 0; dn < DIM; dn++) {
                const int m = 2 * dn + ((c >> dn) & 1);   // the face of cell c on the box boundary
                K[dn] = W(d.K[m]);
                neg |= (d.fly[m] == -1 ? 1 : 0) << dn;
            }
            lpld = lpld_box(T, AUTO, d.slot);
            if (lpld) {
                const R *e = own.eps + pos;
                const W a0 = W(e[0]);
#pragma unroll
                for (int m = 0; m < 2 * DIM; m++) {
                    const W a = W(e[dq[m]]);
                    cm[m] = 2 * a0 * a / (a0 + a);
                    sc = m == 0 ? cm[m] : sc + cm[m];
                }
            }
        }
        for (int rb = 1; rb <= 2 * A.n_base; rb++) {
            W u[6];
#pragma unroll
            for (int dn = 0; dn < DIM; dn++) {
                const W other = xshfl(ph, 1 << dn);   // every lane takes part
                const W ghost = K[dn] + (((neg >> dn) & 1) ? -ph : ph);
                const int bit = (c >> dn) & 1;
#pragma unroll
                for (int high = 0; high < 2; high++) {
                    const int m = 2 * dn + high;
                    if (rb == 1)
                        u[m] = on ? W(ldv(own.phi + pos + dq[m])) : W(0);
                    else
                        u[m] = bit != high ? other : ghost;
                }
            }
            if (DIM == 2) u[4] = u[5] = 0;
            W nv;
            if (lpld) {
                W su = 0;
#pragma unroll
                for (int m = 0; m < 2 * DIM; m++) su = m == 0 ? cm[m] * u[m] : su + cm[m] * u[m];
                nv = (su - W(dx2) * rh) / sc;
            } else {
                nv = gsrb_value<DIM>(u, rh, (const W *)nullptr, 0, S, dx2, false);
            }
            if (odd == (rb & 1)) ph = nv;
        }
        if (on) own.phi[pos] = R(ph);
    }
}

// The base relaxation from the cell list B.base (BOP_BASE when the fast
// path does not apply): the 2 n_base colour passes of gsrb_boxes on the
// lowest level L, the first reading the stored ghosts, the others the
// neighbours the ghost fill after the previous pass would store (cells of
// the other colour, or K + c1 phi at physical boundaries).  Threads t of n;
// sy() orders the passes among them.
template <int DIM, int AUTO, typename R, typename Sy>
AFG_DEV void base_list_run(const DevTree &T, const LevelViewT<R> &own, const BottomLevels &B, const BottomArgs &A,
                           const LevelArgs &L, int t, int n, Sy sy)
{
    const int s = L.nc + 2;
    const int dq[6] = {-1, 1, -s, s, -s * s, s * s};
    auto relax = [&](const BaseCell &d, int rb) {
        const R center = ldv(own.phi + d.pos);
        R u[6];
#pragma unroll
        for (int m = 0; m < 2 * DIM; m++) {
            const int f = d.fly[m];
            if (rb == 1 || f >= 0)
                u[m] = ldv(own.phi + (rb == 1 ? d.pos + dq[m] : f));
            else
                u[m] = R(d.K[m]) + (f == -1 ? -center : center);
        }
        if (DIM == 2) u[4] = u[5] = 0;
        const bool lpld = lpld_box(T, AUTO, d.slot);
        own.phi[d.pos] = gsrb_value<DIM>(u, ldv(own.rhs + d.pos), lpld ? own.eps + d.pos : nullptr, 0, s, L.dx2,
                                         lpld);
    };
    const int n1 = B.base_n[0], n2 = B.base_n[1];
    if (n1 <= n && n2 <= n) {   // one cell of each colour per thread, kept in registers
        BaseCell d1{}, d2{};
        if (t < n1) d1 = B.base[t];
        if (t < n2) d2 = B.base[n1 + t];
        for (int rb = 1; rb <= 2 * A.n_base; rb++) {
            if (rb & 1) {
                if (t < n1) relax(d1, rb);
            } else {
                if (t < n2) relax(d2, rb);
            }
            if (rb < 2 * A.n_base) sy();
        }
    } else {
        for (int rb = 1; rb <= 2 * A.n_base; rb++) {
            const int c0 = rb & 1 ? 0 : n1, nn = rb & 1 ? n1 : n2;
            for (int c = t; c < nn; c += n) relax(B.base[c0 + c], rb);
            if (rb < 2 * A.n_base) sy();
        }
    }
}

// The whole bottom part for thread tid of nt; `lds` is the workgroup's
// dynamic LDS and `sync()` the workgroup barrier; AUTO is the operator
// choice of the options (mg$D_auto_op with eps, or mg$D_box_lpl).  Every thread walks the
// same step sequence.  Steps on levels 0..wave_top run in threads
// [0, wave) alone, one wavefront, and are ordered by `wsync()` among
// themselves; the other threads pass them by and wait at the next workgroup
// barrier.
// TAU: the FP32 stages' update_coarse in FP64 arithmetic (MgVarsT::tau64), a
// compile-time choice so that the other instances hold none of its code.
template <int DIM, int AUTO, typename R, int TAU = 0, typename Sync, typename WaveSync, typename Xshfl>
AFG_DEV void bottom_run(const DevTree &Tg, const MgVarsT<R> &V, const BottomLevels &B,
                        const BottomArgs &A, void *lds, int tid, int nt, int wave, Sync sync,
                        WaveSync wsync, Xshfl xshfl)
{
    constexpr int NNB = 2 * DIM;
    constexpr int BFLY = DIM == 2 ? 1 : 0;   // the relaxation blocks' form (relax_step)
    // topology tables of the bottom slots, in LDS
    char *tb = (char *)lds;
    int32_t *t_nb = (int32_t *)tb;
    int32_t *t_par = t_nb + B.n_slots * NNB;
    int32_t *t_co = t_par + B.n_slots;
    int32_t *t_bct = t_co + B.n_slots;
    int32_t *t_bcv = t_bct + B.n_bc;
    int32_t *t_list = t_bcv + B.n_bc;
    int8_t *t_tag = (int8_t *)(t_list + B.n_par);
    int8_t *t_isp = t_tag + B.n_slots;
    for (int e = tid; e < B.n_slots * NNB; e += nt) t_nb[e] = Tg.nb[e];
    for (int e = tid; e < B.n_slots; e += nt) {
        t_par[e] = Tg.parent[e];
        t_co[e] = Tg.choff[e];
        t_tag[e] = Tg.tag[e];
        t_isp[e] = Tg.is_parent[e];
    }
    for (int e = tid; e < B.n_bc; e += nt) {
        t_bct[e] = Tg.bc_type[e];
        t_bcv[e] = Tg.bc_voff[e];
    }
    for (int li = 0; li <= A.hl; li++)
        for (int e = tid; e < B.par[li].n; e += nt) t_list[B.par_off[li] + e] = B.par[li].slots[e];
    DevTree T = Tg;
    T.nb = t_nb;
    T.parent = t_par;
    T.choff = t_co;
    T.bc_type = t_bct;
    T.bc_voff = t_bcv;
    T.tag = t_tag;
    T.is_parent = t_isp;
    R *data = (R *)(tb + B.table_bytes);

    // the levels' values: one pass over all levels, each element's loads of
    // every variable (and those of a second element) issued before any is
    // stored, so that few memory round trips stand before the first step
    auto copy_in_all = [&] {
        int total = 0;
        for (int li = 0; li <= A.hl; li++) total += B.lv[li].n * B.box_len[li];
        auto locate = [&](int e, int &li, int &r) {
            li = 0;
            r = e;
            while (r >= B.lv[li].n * B.box_len[li]) {
                r -= B.lv[li].n * B.box_len[li];
                li++;
            }
        };
        for (int e = tid; e < total; e += 2 * nt) {
            const int e2 = e + nt;
            int li, r, li2 = 0, r2 = 0;
            locate(e, li, r);
            const bool two = e2 < total;
            if (two) locate(e2, li2, r2);
            const int b = r / B.box_len[li], b2 = two ? r2 / B.box_len[li2] : 0;
            const size_t g = (size_t)(B.lv[li].slot0 + b) * T.stride + (r - b * B.box_len[li]);
            const size_t g2 = two ? (size_t)(B.lv[li2].slot0 + b2) * T.stride + (r2 - b2 * B.box_len[li2]) : g;
            const R p0 = V.phi[g], p1 = V.rhs[g], p2 = V.tmp[g];
            const R q0 = V.phi[g2], q1 = V.rhs[g2], q2 = V.tmp[g2];
            const R p3 = AUTO ? V.eps[g] : R(0), q3 = AUTO ? V.eps[g2] : R(0);
            data[B.lds_off[0][li] + r] = p0;
            data[B.lds_off[1][li] + r] = p1;
            data[B.lds_off[2][li] + r] = p2;
            if (AUTO) data[B.lds_off[3][li] + r] = p3;
            if (two) {
                data[B.lds_off[0][li2] + r2] = q0;
                data[B.lds_off[1][li2] + r2] = q1;
                data[B.lds_off[2][li2] + r2] = q2;
                if (AUTO) data[B.lds_off[3][li2] + r2] = q3;
            }
        }
    };
    auto copy_out = [&](R *dst, int off, int li) {
        const int len = B.box_len[li], n = B.lv[li].n, slot0 = B.lv[li].slot0;
        for (int e = tid; e < n * len; e += nt) {
            const int b = e / len;
            dst[(size_t)(slot0 + b) * T.stride + (e - b * len)] = data[off + e];
        }
    };
    copy_in_all();
    sync();

    // every level of the kernel is in LDS (the host takes only levels that
    // fit), so the views hold LDS addresses only
    auto view = [&](int li) {
        LevelViewT<R> w;
        w.phi = data + B.lds_off[0][li];
        w.rhs = data + B.lds_off[1][li];
        w.tmp = data + B.lds_off[2][li];
        w.eps = AUTO ? data + B.lds_off[3][li] : nullptr;
        w.slot0 = B.lv[li].slot0;
        w.stride = B.box_len[li];
        return w;
    };
    // The steps of one cell count are compiled together (dispatch on the
    // count first, then on the operation): a level's steps then run from one
    // compact range of code, which stays in the instruction cache.
    // One step on threads t of n; `in_wave` picks the barrier that orders the
    // passes of BOP_BASE and the steps of a block.  A relaxation block runs
    // as one step, its level's arguments read once for all its passes and
    // fills.
    auto exec = [&](const BottomStep st, int t, int n, bool in_wave) {
        const int li = st.li;
        const LevelViewT<R> own = view(li), par = view(li > 0 ? li - 1 : 0);
        const bool on_parent = st.op == BOP_UC_PARENT || st.op == BOP_DIFF;
        bottom_nc_dispatch(B.lv[on_parent ? li - 1 : li].nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            switch (st.op) {
            case BOP_BASE: {
                // 2 n_base colour passes of gsrb_boxes on level 0: the first
                // reads the stored ghosts, the others the neighbours the ghost
                // fill after the previous pass would store (cells of the other
                // colour, or K + c1 phi at physical boundaries); then the
                // level's ghost fill
                const LevelArgs L = B.lv[0];
                auto relax = [&] {
                    if constexpr (NC == 2) {
                        if (B.base_fast && wave == 64) {
                            // cells of a box on the lanes of a wavefront, the boxes' groups
                            // over every wavefront of the threads of this step
                            const int lane = t & 63, w = t >> 6, nw = n >> 6;
                            if constexpr (sizeof(R) == 4) {
                                if (A.base64) {
                                    base_wave_run<DIM, AUTO, R, double>(T, own, B, A, L.dx2, L.n, lane, xshfl, w, nw);
                                    return;
                                }
                            }
                            base_wave_run<DIM, AUTO, R, R>(T, own, B, A, L.dx2, L.n, lane, xshfl, w, nw);
                            return;
                        }
                    }
                    base_list_run<DIM, AUTO, R>(T, own, B, A, L, t, n, [&] {
                        if (in_wave)
                            wsync();
                        else
                            sync();
                    });
                };
                relax();
                if (in_wave)
                    wsync();
                else
                    sync();
                const int tf = fill_threads<DIM>(L.n, NC ? NC : L.nc);
                for (int g = t; g < tf; g += n) fill_body<DIM, NC>(T, L, own, par, g);
            } break;
            case BOP_BLOCK: {
                const LevelArgs L = B.lv[li];
                const int tg = gsrb_threads<DIM>(L.n, NC ? NC : L.nc);
                const int tf = fill_threads<DIM>(L.n, NC ? NC : L.nc);
                if (st.lead) {
                    for (int g = t; g < tf; g += n) fill_body<DIM, NC>(T, L, own, par, g);
                    if (st.m > 0) {
                        if (in_wave)
                            wsync();
                        else
                            sync();
                    }
                }
                for (int r = 0; r < st.m; r++) {
                    const BottomStep p = relax_step<BFLY>(li, r, st.rb, st.fly);
                    if (p.op == BOP_GSRB) {
                        for (int g = t; g < tg; g += n)
                            gsrb_body<DIM, NC>(T, L, own, par, AUTO, p.rb, BFLY ? p.fly : 0, g);
                    } else {
                        for (int g = t; g < tf; g += n) fill_body<DIM, NC>(T, L, own, par, g);
                    }
                    if (r + 1 < st.m) {
                        if (in_wave)
                            wsync();
                        else
                            sync();
                    }
                }
            } break;
            case BOP_FILL: {
                const LevelArgs L = B.lv[li];
                const int tot = fill_threads<DIM>(L.n, NC ? NC : L.nc);
                for (int g = t; g < tot; g += n) fill_body<DIM, NC>(T, L, own, par, g);
            } break;
            case BOP_UC_CHILD: {
                const LevelArgs L = B.lv[li];
                const int tot = L.n * ipow((NC ? NC : L.nc) / 2, DIM);
                const int fly = BFLY && A.n_down > 0 ? 1 : 0;
                if constexpr (sizeof(R) == 4 && TAU) {
                    for (int g = t; g < tot; g += n) uc_child_body<DIM, NC, R, double>(T, L, own, par, AUTO, 1, fly, g);
                    break;
                }
                for (int g = t; g < tot; g += n) uc_child_body<DIM, NC>(T, L, own, par, AUTO, 1, fly, g);
            } break;
            case BOP_UC_PARENT: {   // parent part on level li - 1
                const LevelArgs Lp = B.lv[li - 1];
                const LevelViewT<R> ppar = view(li > 1 ? li - 2 : 0);
                const int tot = Lp.n * ipow((NC ? NC : Lp.nc) + 2, DIM);
                if constexpr (!BFLY) {
                    // 3D: the level's ghost fill first, then the parent part
                    // reads the stored ghosts (the same values as computing
                    // them in its stencils, with no divergent ghost code there)
                    const int tf = fill_threads<DIM>(Lp.n, NC ? NC : Lp.nc);
                    for (int g = t; g < tf; g += n) fill_body<DIM, NC>(T, Lp, par, ppar, g);
                    if (in_wave)
                        wsync();
                    else
                        sync();
                }
                if constexpr (sizeof(R) == 4 && TAU) {
                    for (int g = t; g < tot; g += n)
                        uc_parent_body<DIM, NC, R, double>(T, Lp, par, ppar, AUTO, 1, g, BFLY);
                    break;
                }
                for (int g = t; g < tot; g += n) uc_parent_body<DIM, NC>(T, Lp, par, ppar, AUTO, 1, g, BFLY);
            } break;
            case BOP_CORRECT: {
                const LevelArgs L = B.lv[li];
                const int tot = L.n * ipow((NC ? NC : L.nc) / 2, DIM);
                for (int g = t; g < tot; g += n) correct_body<DIM, NC>(T, L, own, par, AUTO, 1, g);
            } break;
            default: {   // BOP_DIFF: parents of level li, on level li - 1
                ListArgs P = B.par[li - 1];
                P.slots = t_list + B.par_off[li - 1];
                const int tot = P.n * ipow((NC ? NC : P.nc) + 2, DIM);
                for (int g = t; g < tot; g += n) diff_body<DIM, NC>(T, P, par, g);
            } break;
            }
        });
    };

    const int W = A.hl < B.wave_top ? A.hl : B.wave_top;
    BottomSeq q = seq_first<BFLY>(A);
    BottomStep st = seq_step<BFLY>(A, q);
    while (q.part < 3) {
        const bool in_wave = st.li <= W;
        if (!in_wave || tid < wave) exec(st, tid, in_wave ? wave : nt, in_wave);
        seq_next<BFLY>(A, q);
        const BottomStep nx = seq_step<BFLY>(A, q);
        if (in_wave && nx.li <= W)
            wsync();
        else
            sync();
        st = nx;
    }

    for (int li = 0; li <= A.hl; li++) {
        copy_out(V.phi, B.lds_off[0][li], li);
        copy_out(V.rhs, B.lds_off[1][li], li);
        copy_out(V.tmp, B.lds_off[2][li], li);
    }
}

// residual on the small levels 0..hl in one launch (set_residual part of
// mg$D_fas_vcycle, m_aX_mg.f90:243-254), on the HBM pools
template <int DIM, typename R>
AFG_DEV void bottom_residual_run(const DevTree &T, const MgVarsT<R> &V, const BottomLevels &B,
                                 const BottomArgs &A, int tid, int nt)
{
    const LevelViewT<R> v = LevelViewT<R>{V.phi, V.rhs, V.tmp, V.eps, 0, T.stride};
    for (int li = 0; li <= A.hl; li++) {
        const LevelArgs &L = B.lv[li];
        nc_dispatch(L.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = L.n * ipow(L.nc, DIM);
            for (int g = tid; g < tot; g += nt) residual_body<DIM, NC>(T, L, v, v, V.op_auto, g);
        });
    }
}


// residual_box (m_aX_mg.f90:545-559) on every cell of levels 0..hl: thread t of
// n takes cells t, t + n, ... of the levels in order (each level a
// contiguous slot range; the operators of the options)
template <int DIM, typename R>
AFG_DEV void residual_levels_run(const DevTree &T, const MgVarsT<R> &V, const BottomLevels &B, int hl, int t,
                                 int n)
{
    const LevelViewT<R> v{V.phi, V.rhs, V.tmp, V.eps, 0, T.stride};
    for (int li = 0; li <= hl; li++) {
        const LevelArgs &L = B.lv[li];
        nc_dispatch(L.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = L.n * ipow(L.nc, DIM);
            for (int g = t; g < tot; g += n) residual_body<DIM, NC>(T, L, v, v, V.op_auto, g);
        });
    }
}

// ---------------------------------------------------------------------------
// The lower part of a V-cycle in one launch of several workgroups (the
// persistent kernel): levels 0..hl (relative to the lowest) with their data
// in the HBM pools, every stage of mg$D_fas_vcycle below and at hl
// (m_aX_mg.f90:218-241) in the order and with the fusions of the per-level
// launches of the host orchestration -- colour passes with face ghosts
// computed from the second pass on, update_coarse as its child and parent
// parts, the base relaxation and its ghost fill, correct_children with the
// correction formed on the fly, the up-sweep's passes and the fill that also
// forms the parents' tmp = phi - tmp -- so the values are those of the
// per-level path bit for bit.  Thread tid of nt in workgroup `rank` of
// n_rank; gsync() orders the stages across all workgroups, wgsync() the
// threads of one workgroup.  With set_residual the residual of levels 0..hl
// is the last stage (only when hl is the V-cycle's highest level: above it,
// the per-level up-sweep still writes tmp on the parents of level hl + 1).
template <int DIM, int AUTO, typename R, typename GSync, typename WSync, typename Xshfl>
AFG_DEV void persist_run(const DevTree &T, const MgVarsT<R> &V, const BottomLevels &B, const BottomArgs &A,
                         int set_residual, int rank, int n_rank, int tid, int nt, int wave, GSync gsync,
                         WSync wgsync, Xshfl xshfl)
{
    const int t = rank * nt + tid, n = n_rank * nt;
    const LevelViewT<R> v{V.phi, V.rhs, V.tmp, V.eps, 0, T.stride};
    // down: relaxation, then update_coarse (child part; parent part with the
    // parent level's ghost fill)
    for (int li = A.hl; li >= 1; li--) {
        const LevelArgs L = B.lv[li];
        bottom_nc_dispatch(L.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int nc = NC ? NC : L.nc;
            const int tot = gsrb_threads<DIM>(L.n, nc);
            for (int rb = 1; rb <= 2 * A.n_down; rb++) {
                for (int g = t; g < tot; g += n) gsrb_body<DIM, NC>(T, L, v, v, AUTO, rb, rb == 1 ? 0 : 1, g);
                gsync();
            }
            const int tc = L.n * ipow(nc / 2, DIM);
            const int fly = A.n_down > 0 ? 1 : 0;
            if constexpr (sizeof(R) == 4) {
                if (A.tau64) {
                    for (int g = t; g < tc; g += n) uc_child_body<DIM, NC, R, double>(T, L, v, v, AUTO, 1, fly, g);
                    return;
                }
            }
            for (int g = t; g < tc; g += n) uc_child_body<DIM, NC>(T, L, v, v, AUTO, 1, fly, g);
        });
        gsync();
        const LevelArgs Lp = B.lv[li - 1];
        bottom_nc_dispatch(Lp.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = Lp.n * ipow((NC ? NC : Lp.nc) + 2, DIM);
            if constexpr (sizeof(R) == 4) {
                if (A.tau64) {
                    for (int g = t; g < tot; g += n) uc_parent_body<DIM, NC, R, double>(T, Lp, v, v, AUTO, 1, g);
                    return;
                }
            }
            for (int g = t; g < tot; g += n) uc_parent_body<DIM, NC>(T, Lp, v, v, AUTO, 1, g);
        });
        gsync();
    }
    // the lowest level: n_base cycles of gsrb_boxes, then its ghost fill
    if (A.n_base > 0) {
        const LevelArgs L = B.lv[0];
        if (B.base_fast && L.nc == 2 && wave == 64) {
            // boxes are independent: every wavefront of every workgroup takes groups of them
            const int lane = tid & 63, w = t >> 6, nw = n >> 6;
            if (sizeof(R) == 4 && A.base64)
                base_wave_run<DIM, AUTO, R, double>(T, v, B, A, L.dx2, L.n, lane, xshfl, w, nw);
            else
                base_wave_run<DIM, AUTO, R, R>(T, v, B, A, L.dx2, L.n, lane, xshfl, w, nw);
        } else if (rank == 0) {
            base_list_run<DIM, AUTO, R>(T, v, B, A, L, tid, nt, wgsync);
        }
        gsync();
        bottom_nc_dispatch(L.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = fill_threads<DIM>(L.n, NC ? NC : L.nc);
            for (int g = t; g < tot; g += n) fill_body<DIM, NC>(T, L, v, v, g);
        });
        gsync();
    }
    // up: correction, relaxation, then the ghost fill with tmp = phi - tmp
    // on the parents of the level below
    for (int li = 1; li <= A.hl; li++) {
        const LevelArgs L = B.lv[li];
        const ListArgs P = B.par[li - 1];
        bottom_nc_dispatch(L.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int nc = NC ? NC : L.nc;
            const int tc = L.n * ipow(nc / 2, DIM);
            for (int g = t; g < tc; g += n) correct_body<DIM, NC>(T, L, v, v, AUTO, 1, g);
            gsync();
            const int tot = gsrb_threads<DIM>(L.n, nc);
            for (int rb = 1; rb <= 2 * A.n_up; rb++) {
                for (int g = t; g < tot; g += n) gsrb_body<DIM, NC>(T, L, v, v, AUTO, rb, 1, g);
                gsync();
            }
            const int nf = fill_threads<DIM>(L.n, nc);
            const int tf = nf + P.n * ipow(P.nc + 2, DIM);
            for (int g = t; g < tf; g += n) {
                if (g < nf)
                    fill_body<DIM, NC>(T, L, v, v, g);
                else
                    diff_body<DIM, 0>(T, P, v, g - nf);
            }
        });
        if (li < A.hl || set_residual) gsync();
    }
    if (set_residual) residual_levels_run<DIM, R>(T, V, B, A.hl, t, n);
}


}  // namespace afg
