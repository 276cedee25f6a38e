// afg_kernels.hip -- gfx950 kernels of the afivo_amd FAS multigrid path.
//
// Every kernel restates one afivo routine over all boxes of a level (the
// reference runs these as OpenMP loops over boxes).  Floating-point
// expressions keep the reference's evaluation order and the library is built
// with -ffp-contract=off, so results are bit-identical to the reference.
//
// Box arrays: var + slot * stride holds cc(0:nc+1, 0:nc+1[, 0:nc+1]) with x
// fastest; s = nc + 2 is the row pitch.  NC is the compile-time cell count
// per dimension (0: taken from the launch arguments).

#include "afg_internal.h"

namespace afg {

namespace {

constexpr int BLOCK = 256;

__host__ __device__ constexpr int ipow(int n, int d) { return d == 2 ? n * n : n * n * n; }

inline int n_blocks(long long threads, int block = BLOCK)
{
    return (int)((threads + block - 1) / block);
}

// ---------------------------------------------------------------------------
// ghost value of face f next to interior cell (i, j, k) of the box in `slot`
//
// The one expression used both by the ghost-fill kernel and, on the fly, by
// the colour passes.  Restates a$D_gc_box (m_aX_gc.f90:50-69):
//   neighbour at the same level  -> sides_from_nb  (m_aX_gc.f90:381-415)
//   refinement boundary          -> mg$D_sides_rb  (m_aX_mg.f90:262-350) after
//                                   a$D_gc_prolong0 (m_aX_gc.f90:135-151,
//                                   m_aX_prolong.f90:39-108)
//   physical boundary            -> captured sides_bc value b, then bc_to_gc
//                                   (m_aX_gc.f90:71-132): c2*b + c1*interior
// `center` is the interior cell next to the ghost, `box` its box array.
template <int DIM>
__device__ __forceinline__ double ghost_value(const DevTree &T, const double *__restrict__ pool,
                                              const double *__restrict__ box, int slot, int f,
                                              int nc, int pnc, double dr, int i, int j, int k,
                                              double center)
{
    const int s = nc + 2;
    const int nbv = T.nb[slot * (2 * DIM) + f];
    const int dn = f >> 1, high = f & 1;
    if (nbv >= 0) {
        const int m = high ? 1 : nc;
        const int a0 = dn == 0 ? m : i, a1 = dn == 1 ? m : j, a2 = dn == 2 ? m : k;
        return pool[(size_t)nbv * T.stride + a0 + s * (a1 + s * a2)];
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
        const double P = pool[(size_t)p * T.stride + pi + ps * (pj + ps * pk)];
        const int inward = high ? -1 : 1;
        if (DIM == 2) {
            const int di = dn == 0 ? inward : -1 + 2 * (i & 1);
            const int dj = dn == 1 ? inward : -1 + 2 * (j & 1);
            const double a = box[(i + di) + s * j];
            const double b = box[i + s * (j + dj)];
            return 0.5 * P + center - 0.25 * (a + b);
        } else {
            const int d0 = dn == 0 ? inward : -1 + 2 * (i & 1);
            const int d1 = dn == 1 ? inward : -1 + 2 * (j & 1);
            const int d2 = dn == 2 ? inward : -1 + 2 * (k & 1);
            const double d = box[(i + d0) + s * ((j + d1) + s * (k + d2))];
            return 0.5 * P + 0.75 * center - 0.25 * d;
        }
    }
    const int rec = NB_BC0 - nbv;
    const int voff = T.bc_voff[rec];
    double b = 0.0;
    if (voff >= 0) {
        const int t1 = dn == 0 ? j : i;
        const int t2 = dn == 2 ? j : k;
        b = T.bc_vals[voff + (t1 - 1) + (DIM == 3 ? nc * (t2 - 1) : 0)];
    }
    double c1, c2;
    if (T.bc_type[rec] == -1) {   // a5_bc_dirichlet
        c1 = -1;
        c2 = 2;
    } else {                      // a5_bc_neumann
        c1 = 1;
        c2 = dr * (high ? 1.0 : -1.0);
    }
    return c2 * b + c1 * center;
}

// ---------------------------------------------------------------------------
// ghost fill of a level: fill_gc_phi (m_aX_mg.f90:352-365)
// one thread per face ghost cell: n * 2*DIM * nc^(DIM-1)

template <int DIM, int NC>
__device__ __forceinline__ void fill_gc_body(const DevTree &T, const LevelArgs &L, double *var,
                                             int gid)
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
    double *box = var + (size_t)slot * T.stride;
    const double center = box[i + s * (j + s * k)];
    const double g = ghost_value<DIM>(T, var, box, slot, f, nc, L.pnc, L.dr, i, j, k, center);
    const int gi = dn == 0 ? (high ? nc + 1 : 0) : i;
    const int gj = dn == 1 ? (high ? nc + 1 : 0) : j;
    const int gk = dn == 2 ? (high ? nc + 1 : 0) : k;
    box[gi + s * (gj + s * gk)] = g;
}

template <int DIM>
__host__ __device__ inline int fill_gc_threads(int n, int nc)
{
    return n * 2 * DIM * (DIM == 2 ? nc : nc * nc);
}

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_fill_gc(DevTree T, LevelArgs L, double *var)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < fill_gc_threads<DIM>(L.n, NC ? NC : L.nc)) fill_gc_body<DIM, NC>(T, L, var, gid);
}

// ---------------------------------------------------------------------------
// one colour pass of GSRB over a level: the body of gsrb_boxes
// (m_aX_mg.f90:404-411) with mg$D_box_gsrb_lpl (m_aX_mg.f90:718-760) or, for
// tagged boxes under the auto operators, mg$D_box_gsrb_lpld (:810-863).
//
// With stored = 1 the face neighbours are read from the ghost layer in
// memory, as the reference does.  With stored = 0 they are computed on the
// fly with ghost_value from the state at the start of the pass, which equals
// what the reference's ghost fill after the previous pass stores: a face
// ghost feeds only the interior cell next to it, and every value it depends
// on (that cell, the neighbour's mirrored cell, the parent) is untouched by
// this pass because nc is even.  One thread per cell of colour rb.

template <int DIM, int NC>
__device__ __forceinline__ void gsrb_body(const DevTree &T, const LevelArgs &L, const MgVars &V,
                                          int rb, int stored, int gid)
{
    const int nc = NC ? NC : L.nc;
    const int hn = nc >> 1;
    const int half = DIM == 2 ? nc * hn : nc * nc * hn;
    const int b = gid / half;
    const int c = gid - b * half;
    const int r = c / hn;
    const int p = c - r * hn;
    const int j = DIM == 2 ? r + 1 : r % nc + 1;
    const int k = DIM == 2 ? 0 : r / nc + 1;
    const int i = 2 - ((rb ^ (DIM == 2 ? j : k + j)) & 1) + 2 * p;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    double *box = V.phi + (size_t)slot * T.stride;
    const int q = i + s * (j + s * k);
    const double center = box[q];
    const double xm = (i > 1 || stored) ? box[q - 1]
        : ghost_value<DIM>(T, V.phi, box, slot, 0, nc, L.pnc, L.dr, i, j, k, center);
    const double xp = (i < nc || stored) ? box[q + 1]
        : ghost_value<DIM>(T, V.phi, box, slot, 1, nc, L.pnc, L.dr, i, j, k, center);
    const double ym = (j > 1 || stored) ? box[q - s]
        : ghost_value<DIM>(T, V.phi, box, slot, 2, nc, L.pnc, L.dr, i, j, k, center);
    const double yp = (j < nc || stored) ? box[q + s]
        : ghost_value<DIM>(T, V.phi, box, slot, 3, nc, L.pnc, L.dr, i, j, k, center);
    double zm = 0, zp = 0;
    if (DIM == 3) {
        zm = (k > 1 || stored) ? box[q - s * s]
            : ghost_value<DIM>(T, V.phi, box, slot, 4, nc, L.pnc, L.dr, i, j, k, center);
        zp = (k < nc || stored) ? box[q + s * s]
            : ghost_value<DIM>(T, V.phi, box, slot, 5, nc, L.pnc, L.dr, i, j, k, center);
    }
    const double rhs = V.rhs[(size_t)slot * T.stride + q];
    if (V.op_auto && T.tag[slot] != TAG_NORMAL) {
        // mg$D_box_gsrb_lpld: c = 2 a0 a / (a0 + a); phi = (sum(c u) - dx2 rhs) / sum(c)
        const double *e = V.eps + (size_t)slot * T.stride;
        const double a0 = e[q];
        const double u[6] = {xm, xp, ym, yp, zm, zp};
        const int dq[6] = {-1, 1, -s, s, -s * s, s * s};
        double su = 0, sc = 0;
#pragma unroll
        for (int m = 0; m < 2 * DIM; m++) {
            const double a = e[q + dq[m]];
            const double cm = 2 * a0 * a / (a0 + a);
            if (m == 0) {
                su = cm * u[m];
                sc = cm;
            } else {
                su += cm * u[m];
                sc += cm;
            }
        }
        box[q] = (su - L.dx2 * rhs) / sc;
    } else if (DIM == 2) {
        box[q] = 0.25 * (xp + xm + yp + ym - L.dx2 * rhs);
    } else {
        constexpr double sixth = 1 / 6.0;
        box[q] = sixth * (xp + xm + yp + ym + zp + zm - L.dx2 * rhs);
    }
}

template <int DIM>
__host__ __device__ inline int gsrb_threads(int n, int nc)
{
    return n * ipow(nc, DIM) / 2;
}

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_gsrb_pass(DevTree T, LevelArgs L, MgVars V, int rb,
                                                     int stored)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < gsrb_threads<DIM>(L.n, NC ? NC : L.nc)) gsrb_body<DIM, NC>(T, L, V, rb, stored, gid);
}

// ---------------------------------------------------------------------------
// operator A phi at interior cell q of a box: mg$D_box_lpl (m_aX_mg.f90:763-797)
// or mg$D_box_lpld (m_aX_mg.f90:866-916) for tagged boxes.

template <int DIM>
__device__ __forceinline__ double box_op_at(const double *__restrict__ ph,
                                            const double *__restrict__ e, int q, int s,
                                            double inv_dr_sq, bool lpld)
{
    if (lpld) {
        const int dq[6] = {-1, 1, -s, s, -s * s, s * s};
        const double u0 = ph[q], a0 = e[q];
        double sum = 0;
#pragma unroll
        for (int m = 0; m < 2 * DIM; m++) {
            const double a = e[q + dq[m]];
            const double t = a0 * a / (a0 + a) * (ph[q + dq[m]] - u0);
            sum = m == 0 ? t : sum + t;
        }
        return inv_dr_sq * 2 * sum;
    }
    if (DIM == 2)
        return inv_dr_sq * (ph[q - 1] + ph[q + 1] + ph[q - s] + ph[q + s] - 4 * ph[q]);
    return inv_dr_sq * (ph[q - 1] + ph[q + 1] + ph[q - s] + ph[q + s] + ph[q - s * s] +
                        ph[q + s * s] - 6 * ph[q]);
}

// residual_box (m_aX_mg.f90:545-559) over a level: tmp = rhs - A phi

template <int DIM, int NC>
__device__ __forceinline__ void residual_body(const DevTree &T, const LevelArgs &L, const MgVars &V,
                                              int gid)
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
    const size_t base = (size_t)slot * T.stride;
    const int q = i + s * (j + s * k);
    const bool lpld = V.op_auto && T.tag[slot] != TAG_NORMAL;
    const double a = box_op_at<DIM>(V.phi + base, lpld ? V.eps + base : nullptr, q, s,
                                    L.inv_dr_sq, lpld);
    V.tmp[base + q] = V.rhs[base + q] - a;
}

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_residual(DevTree T, LevelArgs L, MgVars V)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow(NC ? NC : L.nc, DIM)) residual_body<DIM, NC>(T, L, V, gid);
}

// ---------------------------------------------------------------------------
// update_coarse, child part (m_aX_mg.f90:444-465), fused: per coarse cell of a
// child's octant, compute the residual of its 2^DIM fine cells, restrict it
// into the parent's tmp and restrict phi into the parent's phi
// (a$D_restrict_box, m_aX_restrict.f90:67-132).  The reference stores the
// residual in the child's tmp and restores tmp afterwards; with
// keep_child_tmp = 0 (set_coarse_phi_rhs, m_aX_mg.f90:495-504) the residual
// stays in the child's tmp.

template <int DIM, int NC>
__device__ __forceinline__ void update_coarse_child_body(const DevTree &T, const LevelArgs &L,
                                                         const MgVars &V, int keep_child_tmp,
                                                         int gid)
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
    const size_t base = (size_t)slot * T.stride;
    const bool lpld = V.op_auto && T.tag[slot] != TAG_NORMAL;
    const double *ph = V.phi + base;
    const double *e = lpld ? V.eps + base : nullptr;
    const double *rh = V.rhs + base;
    double sr = 0, sp = 0;
#pragma unroll
    for (int m = 0; m < (1 << DIM); m++) {
        const int fi = 2 * ii - 1 + (m & 1);
        const int fj = 2 * jj - 1 + ((m >> 1) & 1);
        const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
        const int q = fi + s * (fj + s * fk);
        const double res = rh[q] - box_op_at<DIM>(ph, e, q, s, L.inv_dr_sq, lpld);
        if (!keep_child_tmp) V.tmp[base + q] = res;
        sr = m == 0 ? res : sr + res;
        sp = m == 0 ? ph[q] : sp + ph[q];
    }
    const int p = T.parent[slot];
    const int co = T.choff[slot];
    const int ps = L.pnc + 2;
    const int pq = ((co & 1023) + ii) +
                   ps * ((((co >> 10) & 1023) + jj) + ps * (DIM == 3 ? ((co >> 20) & 1023) + kk : 0));
    const size_t pb = (size_t)p * T.stride;
    constexpr double w = DIM == 2 ? 0.25 : 0.125;
    V.tmp[pb + pq] = w * sr;
    V.phi[pb + pq] = w * sp;
}

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_update_coarse_child(DevTree T, LevelArgs L, MgVars V,
                                                               int keep_child_tmp)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) / 2, DIM))
        update_coarse_child_body<DIM, NC>(T, L, V, keep_child_tmp, gid);
}

// update_coarse, parent part (m_aX_mg.f90:472-479): rhs = A phi + tmp on the
// interior, rhs = rhs + tmp on ghost and corner cells (a$D_box_add_cc over the
// whole array), then tmp = phi (a$D_box_copy_cc) if copy_tmp.
// One thread per cell of the whole (nc+2)^DIM array of each parent.

template <int DIM, int NC>
__device__ __forceinline__ void update_coarse_parent_body(const DevTree &T, const ListArgs &P,
                                                          const MgVars &V, int copy_tmp, int gid)
{
    const int nc = NC ? NC : P.nc;
    const int s = nc + 2;
    const int per_box = ipow(s, DIM);
    const int b = gid / per_box;
    const int q = gid - b * per_box;
    const int i = q % s;
    const int j = (q / s) % s;
    const int k = DIM == 3 ? q / (s * s) : 1;
    const int slot = P.slots[b];
    const size_t base = (size_t)slot * T.stride;
    const bool interior = i >= 1 && i <= nc && j >= 1 && j <= nc && k >= 1 && k <= nc;
    const double tm = V.tmp[base + q];
    double r;
    if (interior) {
        const bool lpld = V.op_auto && T.tag[slot] != TAG_NORMAL;
        r = box_op_at<DIM>(V.phi + base, lpld ? V.eps + base : nullptr, q, s, P.inv_dr_sq, lpld);
    } else {
        r = V.rhs[base + q];
    }
    V.rhs[base + q] = r + tm;
    if (copy_tmp) V.tmp[base + q] = V.phi[base + q];
}

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_update_coarse_parent(DevTree T, ListArgs P, MgVars V,
                                                                int copy_tmp)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < P.n * ipow((NC ? NC : P.nc) + 2, DIM))
        update_coarse_parent_body<DIM, NC>(T, P, V, copy_tmp, gid);
}

// correct_children, parent part (m_aX_mg.f90:380-381): tmp = phi - tmp over
// the whole array.

template <int DIM, int NC>
__device__ __forceinline__ void diff_body(const DevTree &T, const ListArgs &P, const MgVars &V, int gid)
{
    const int nc = NC ? NC : P.nc;
    const int per_box = ipow(nc + 2, DIM);
    const int b = gid / per_box;
    const int q = gid - b * per_box;
    const size_t base = (size_t)P.slots[b] * T.stride;
    V.tmp[base + q] = V.phi[base + q] - V.tmp[base + q];
}

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_diff_phi_tmp(DevTree T, ListArgs P, MgVars V)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < P.n * ipow((NC ? NC : P.nc) + 2, DIM)) diff_body<DIM, NC>(T, P, V, gid);
}

// Quadratic prolongation of the coarse cell (ic, jc, kc) of parent array p:
// the stencil of a$D_prolong2 (m_aX_prolong.f90:281-343).  Returns f0, the
// first differences and f2; the child value of octant (sx, sy, sz) is then
// (((f0 -/+ fx) -/+ fy) -/+ fz) + f2.
template <int DIM>
struct Prolong2 {
    double f0, fx, fy, fz, f2;
    __device__ __forceinline__ Prolong2(const double *__restrict__ p, int pq, int ps)
    {
        f0 = p[pq];
        fx = 0.125 * (p[pq + 1] - p[pq - 1]);
        fy = 0.125 * (p[pq + ps] - p[pq - ps]);
        const double fxx = 0.03125 * (p[pq - 1] - 2 * f0 + p[pq + 1]);
        const double fyy = 0.03125 * (p[pq - ps] - 2 * f0 + p[pq + ps]);
        if (DIM == 3) {
            fz = 0.125 * (p[pq + ps * ps] - p[pq - ps * ps]);
            const double fzz = 0.03125 * (p[pq - ps * ps] - 2 * f0 + p[pq + ps * ps]);
            f2 = fxx + fyy + fzz;
        } else {
            fz = 0;
            f2 = fxx + fyy;
        }
    }
    __device__ __forceinline__ double at(int m) const
    {
        double v = (m & 1) ? f0 + fx : f0 - fx;
        v = ((m >> 1) & 1) ? v + fy : v - fy;
        if (DIM == 3) v = ((m >> 2) & 1) ? v + fz : v - fz;
        return v + f2;
    }
};

// correct_children, child part: mg%box_corr for each child.  For untagged
// boxes mg$D_box_corr_lpl = a$D_prolong2(add) of the parent's tmp into phi
// (m_aX_mg.f90:708-715); under the auto operators tagged boxes use
// mg$D_box_corr_lpld (m_aX_mg.f90:918-985).  One thread per coarse cell of
// the child's octant, updating its 2^DIM fine cells.

template <int DIM, int NC>
__device__ __forceinline__ void correct_body(const DevTree &T, const LevelArgs &L, const MgVars &V,
                                             int gid)
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
    const size_t pb = (size_t)p * T.stride;
    double *cph = V.phi + (size_t)slot * T.stride;
    if (V.op_auto && T.tag[slot] != TAG_NORMAL) {
        const double *pc = V.tmp + pb;
        const double *pe = V.eps + pb;
        constexpr double third = 1 / 3.0;
#pragma unroll
        for (int m = 0; m < (1 << DIM); m++) {
            const int sx = m & 1, sy = (m >> 1) & 1, sz = (m >> 2) & 1;
            const int fi = 2 * ii - 1 + sx, fj = 2 * jj - 1 + sy;
            const int fk = DIM == 3 ? 2 * kk - 1 + sz : 0;
            // closest coarse cell is (ic, jc, kc); one-but-closest offsets:
            // fine index odd -> -1, even -> +1 (m_aX_mg.f90:944-958)
            const int qn[3] = {pq + (sx ? 1 : -1), pq + (sy ? ps : -ps),
                               pq + (sz ? ps * ps : -ps * ps)};
            const double u0 = pc[pq], a0 = pe[pq];
            double *t = &cph[fi + s * (fj + s * fk)];
            if (DIM == 2) {
                double sum = (a0 * u0 + pe[qn[0]] * pc[qn[0]]) / (a0 + pe[qn[0]]);
                sum += (a0 * u0 + pe[qn[1]] * pc[qn[1]]) / (a0 + pe[qn[1]]);
                *t = *t + 0.5 * sum;
            } else {
                double sum = 0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double v = (a0 * u0 + pe[qn[d]] * (1.5 * pc[qn[d]] - 0.5 * u0)) /
                                     (a0 + pe[qn[d]]);
                    sum = d == 0 ? v : sum + v;
                }
                *t = *t + third * sum;
            }
        }
    } else {
        const Prolong2<DIM> P2(V.tmp + pb, pq, ps);
#pragma unroll
        for (int m = 0; m < (1 << DIM); m++) {
            const int fi = 2 * ii - 1 + (m & 1), fj = 2 * jj - 1 + ((m >> 1) & 1);
            const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
            double *t = &cph[fi + s * (fj + s * fk)];
            *t = P2.at(m) + *t;
        }
    }
}

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_correct(DevTree T, LevelArgs L, MgVars V)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) / 2, DIM)) correct_body<DIM, NC>(T, L, V, gid);
}

// a$D_restrict_box (m_aX_restrict.f90:67-132) of every box of a level into
// its parent: one thread per coarse cell.

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_restrict(DevTree T, LevelArgs L, const double *from,
                                                    double *to)
{
    const int nc = NC ? NC : L.nc;
    const int hn = nc >> 1;
    const int per_box = ipow(hn, DIM);
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= L.n * per_box) return;
    const int b = gid / per_box;
    const int c = gid - b * per_box;
    const int ii = c % hn + 1;
    const int jj = (c / hn) % hn + 1;
    const int kk = DIM == 3 ? c / (hn * hn) + 1 : 0;
    const int slot = L.slot0 + b;
    const int s = nc + 2;
    const double *cb = from + (size_t)slot * T.stride;
    double sum = 0;
#pragma unroll
    for (int m = 0; m < (1 << DIM); m++) {
        const int fi = 2 * ii - 1 + (m & 1), fj = 2 * jj - 1 + ((m >> 1) & 1);
        const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
        const double v = cb[fi + s * (fj + s * fk)];
        sum = m == 0 ? v : sum + v;
    }
    const int p = T.parent[slot];
    const int co = T.choff[slot];
    const int ps = L.pnc + 2;
    const int pq = ((co & 1023) + ii) +
                   ps * ((((co >> 10) & 1023) + jj) + ps * (DIM == 3 ? ((co >> 20) & 1023) + kk : 0));
    constexpr double w = DIM == 2 ? 0.25 : 0.125;
    to[(size_t)p * T.stride + pq] = w * sum;
}

// a$D_prolong2 (m_aX_prolong.f90:243-348) from the parent into every box of
// a level; add = 0 overwrites the interior (the reference zeroes it first and
// adds, so the stored value is expr + 0).

template <int DIM, int NC>
__global__ __launch_bounds__(BLOCK) void k_prolong2(DevTree T, LevelArgs L, const double *from,
                                                    double *to, int add)
{
    const int nc = NC ? NC : L.nc;
    const int hn = nc >> 1;
    const int per_box = ipow(hn, DIM);
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= L.n * per_box) return;
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
    const Prolong2<DIM> P2(from + (size_t)p * T.stride, pq, ps);
    double *cb = to + (size_t)slot * T.stride;
#pragma unroll
    for (int m = 0; m < (1 << DIM); m++) {
        const int fi = 2 * ii - 1 + (m & 1), fj = 2 * jj - 1 + ((m >> 1) & 1);
        const int fk = DIM == 3 ? 2 * kk - 1 + ((m >> 2) & 1) : 0;
        double *t = &cb[fi + s * (fj + s * fk)];
        const double old = add ? *t : 0.0;
        *t = P2.at(m) + old;
    }
}

// whole-array fill and copy over a contiguous slot range
// (a$D_box_clear_cc, m_aX_utils.f90:262-270; a$D_box_copy_cc, :331-339)

__global__ __launch_bounds__(BLOCK) void k_fill_value(int stride, int slot0, int n, int len,
                                                      double *var, double value)
{
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)n * len) return;
    const int b = (int)(gid / len);
    const int q = (int)(gid - (long long)b * len);
    var[(size_t)(slot0 + b) * stride + q] = value;
}

__global__ __launch_bounds__(BLOCK) void k_copy_boxes(int stride, int slot0, int n, int len,
                                                      const double *from, double *to)
{
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)n * len) return;
    const int b = (int)(gid / len);
    const int q = (int)(gid - (long long)b * len);
    const size_t o = (size_t)(slot0 + b) * stride + q;
    to[o] = from[o];
}

// ---------------------------------------------------------------------------
// reductions (wave64 shuffles, then LDS across the waves of a block)

__device__ __forceinline__ double pick(double a, double b, int want_max)
{
    return want_max ? (b > a ? b : a) : (b < a ? b : a);
}

__device__ __forceinline__ double block_extreme(double v, int want_max)
{
    __shared__ double part[BLOCK / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = pick(v, __shfl_xor(v, o, 64), want_max);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) v = pick(v, part[k], want_max);
    }
    return v;
}

// mg$D_set_box_tag (m_aX_mg.f90:667-706), eps part: min and max of eps over
// the whole array, one block per box, only for boxes whose tag is unset.
__global__ __launch_bounds__(BLOCK) void k_set_tags(DevTree T, int slot0, int len,
                                                    const double *eps)
{
    const int slot = slot0 + blockIdx.x;
    if (T.tag[slot] != TAG_UNSET) return;
    const double *e = eps + (size_t)slot * T.stride;
    double lo = e[0], hi = e[0];
    for (int q = threadIdx.x; q < len; q += blockDim.x) {
        lo = e[q] < lo ? e[q] : lo;
        hi = e[q] > hi ? e[q] : hi;
    }
    lo = block_extreme(lo, 0);
    __syncthreads();
    hi = block_extreme(hi, 1);
    if (threadIdx.x == 0) {
        int8_t tag = TAG_NORMAL;
        if (hi > lo)
            tag = TAG_VEPS;
        else if (lo < 1 || lo > 1)
            tag = TAG_CEPS;
        T.tag[slot] = tag;
    }
}

// a$D_tree_max_cc / a$D_tree_min_cc (m_aX_utils.f90:468-571) over a slot list:
// one partial per block.
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_extreme_list(DevTree T, ListArgs P, const double *var,
                                                        int want_max, double *partial)
{
    const int nc = P.nc;
    const int per_box = ipow(nc, DIM);
    const long long total = (long long)P.n * per_box;
    double v = want_max ? -1.7976931348623157e308 : 1.7976931348623157e308;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(g / per_box);
        const int c = (int)(g - (long long)b * per_box);
        const int i = c % nc + 1;
        const int j = (c / nc) % nc + 1;
        const int k = DIM == 3 ? c / (nc * nc) + 1 : 0;
        const int s = nc + 2;
        v = pick(v, var[(size_t)P.slots[b] * T.stride + i + s * (j + s * k)], want_max);
    }
    v = block_extreme(v, want_max);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ __launch_bounds__(BLOCK) void k_reduce_partials(const double *partial, int n,
                                                           int want_max, double *out)
{
    double v = want_max ? -1.7976931348623157e308 : 1.7976931348623157e308;
    for (int q = threadIdx.x; q < n; q += blockDim.x) v = pick(v, partial[q], want_max);
    v = block_extreme(v, want_max);
    if (threadIdx.x == 0) out[0] = v;
}

// ---------------------------------------------------------------------------
// bottom of the V-cycle in one workgroup
//
// Runs, for levels hl .. lowest, exactly the stage sequence of the per-level
// path (gsrb_boxes, update_coarse, the base relaxation, correct_children,
// fill_gc_phi; m_aX_mg.f90:218-241) with the same thread bodies, separated by
// workgroup barriers instead of kernel boundaries.

template <int N>
struct IC {
    static constexpr int value = N;
};

template <typename F>
__device__ __forceinline__ void nc_dispatch(int nc, F &&f)
{
    switch (nc) {
    case 2: f(IC<2>()); break;
    case 4: f(IC<4>()); break;
    case 8: f(IC<8>()); break;
    case 16: f(IC<16>()); break;
    default: f(IC<0>()); break;
    }
}

template <int DIM>
__global__ __launch_bounds__(BOTTOM_BLOCK) void k_bottom(DevTree T, MgVars V, BottomArgs A)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const BottomLevels &B = *A.levels;

    auto fill = [&](int li) {
        const LevelArgs &L = B.lv[li];
        nc_dispatch(L.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = fill_gc_threads<DIM>(L.n, L.nc);
            for (int g = tid; g < tot; g += nt) fill_gc_body<DIM, NC>(T, L, V.phi, g);
        });
        __syncthreads();
    };
    auto gsrb = [&](int li, int n_cycle) {
        const LevelArgs &L = B.lv[li];
        for (int rb = 1; rb <= 2 * n_cycle; rb++) {
            nc_dispatch(L.nc, [&](auto ncv) {
                constexpr int NC = decltype(ncv)::value;
                const int tot = gsrb_threads<DIM>(L.n, L.nc);
                for (int g = tid; g < tot; g += nt) gsrb_body<DIM, NC>(T, L, V, rb, rb == 1, g);
            });
            __syncthreads();
        }
        if (n_cycle > 0) fill(li);
    };
    auto update_coarse = [&](int li) {
        const LevelArgs &L = B.lv[li];
        nc_dispatch(L.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = L.n * ipow(L.nc / 2, DIM);
            for (int g = tid; g < tot; g += nt) update_coarse_child_body<DIM, NC>(T, L, V, 1, g);
        });
        __syncthreads();
        fill(li - 1);
        const ListArgs &P = B.par[li - 1];
        nc_dispatch(P.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = P.n * ipow(P.nc + 2, DIM);
            for (int g = tid; g < tot; g += nt) update_coarse_parent_body<DIM, NC>(T, P, V, 1, g);
        });
        __syncthreads();
    };
    auto correct_children = [&](int li) {
        const ListArgs &P = B.par[li - 1];
        nc_dispatch(P.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = P.n * ipow(P.nc + 2, DIM);
            for (int g = tid; g < tot; g += nt) diff_body<DIM, NC>(T, P, V, g);
        });
        __syncthreads();
        const LevelArgs &L = B.lv[li];
        nc_dispatch(L.nc, [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            const int tot = L.n * ipow(L.nc / 2, DIM);
            for (int g = tid; g < tot; g += nt) correct_body<DIM, NC>(T, L, V, g);
        });
        __syncthreads();
    };

    if (A.mode == 1) {
        for (int li = 0; li <= A.hl; li++) {
            const LevelArgs &L = B.lv[li];
            nc_dispatch(L.nc, [&](auto ncv) {
                constexpr int NC = decltype(ncv)::value;
                const int tot = L.n * ipow(L.nc, DIM);
                for (int g = tid; g < tot; g += nt) residual_body<DIM, NC>(T, L, V, g);
            });
        }
        return;
    }
    for (int li = A.hl; li >= 1; li--) {
        gsrb(li, A.n_down);
        update_coarse(li);
    }
    gsrb(0, A.n_base);
    for (int li = 1; li <= A.hl; li++) {
        correct_children(li);
        fill(li);
        gsrb(li, A.n_up);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers: dispatch the run-time cell count to a compiled instance

#define AFG_DISPATCH(dim, nc, KERNEL, GRID, ...)                                   \
    do {                                                                           \
        if ((dim) == 2) {                                                          \
            switch (nc) {                                                          \
            case 2: KERNEL<2, 2><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 4: KERNEL<2, 4><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 8: KERNEL<2, 8><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 16: KERNEL<2, 16><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;     \
            default: KERNEL<2, 0><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;      \
            }                                                                      \
        } else {                                                                   \
            switch (nc) {                                                          \
            case 2: KERNEL<3, 2><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 4: KERNEL<3, 4><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 8: KERNEL<3, 8><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 16: KERNEL<3, 16><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;     \
            default: KERNEL<3, 0><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;      \
            }                                                                      \
        }                                                                          \
    } while (0)

void launch_fill_gc(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, double *var)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * 2 * dim * (dim == 2 ? L.nc : L.nc * L.nc);
    AFG_DISPATCH(dim, L.nc, k_fill_gc, n_blocks(th), T, L, var);
}

void launch_gsrb_pass(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const MgVars &V, int rb, int stored_ghosts)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc, dim) / 2;
    AFG_DISPATCH(dim, L.nc, k_gsrb_pass, n_blocks(th), T, L, V, rb, stored_ghosts);
}

void launch_residual(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const MgVars &V)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc, dim);
    AFG_DISPATCH(dim, L.nc, k_residual, n_blocks(th), T, L, V);
}

void launch_update_coarse_child(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                                const MgVars &V, int keep_child_tmp)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc / 2, dim);
    AFG_DISPATCH(dim, L.nc, k_update_coarse_child, n_blocks(th), T, L, V, keep_child_tmp);
}

void launch_update_coarse_parent(int dim, hipStream_t s, const DevTree &T, const ListArgs &P,
                                 const MgVars &V, int copy_tmp)
{
    if (P.n <= 0) return;
    const long long th = (long long)P.n * ipow(P.nc + 2, dim);
    AFG_DISPATCH(dim, P.nc, k_update_coarse_parent, n_blocks(th), T, P, V, copy_tmp);
}

void launch_diff_phi_tmp(int dim, hipStream_t s, const DevTree &T, const ListArgs &P,
                         const MgVars &V)
{
    if (P.n <= 0) return;
    const long long th = (long long)P.n * ipow(P.nc + 2, dim);
    AFG_DISPATCH(dim, P.nc, k_diff_phi_tmp, n_blocks(th), T, P, V);
}

void launch_correct(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                    const MgVars &V)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc / 2, dim);
    AFG_DISPATCH(dim, L.nc, k_correct, n_blocks(th), T, L, V);
}

void launch_restrict(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const double *from, double *to)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc / 2, dim);
    AFG_DISPATCH(dim, L.nc, k_restrict, n_blocks(th), T, L, from, to);
}

void launch_prolong2(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const double *from, double *to, int add)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc / 2, dim);
    AFG_DISPATCH(dim, L.nc, k_prolong2, n_blocks(th), T, L, from, to, add);
}

void launch_fill_value(int dim, hipStream_t s, const DevTree &T, int slot0, int n_slots, int nc,
                       double *var, double value)
{
    if (n_slots <= 0) return;
    const int len = ipow(nc + 2, dim);
    k_fill_value<<<n_blocks((long long)n_slots * len), BLOCK, 0, s>>>(T.stride, slot0, n_slots,
                                                                      len, var, value);
}

void launch_copy_boxes(int dim, hipStream_t s, const DevTree &T, int slot0, int n_slots, int nc,
                       const double *from, double *to)
{
    if (n_slots <= 0) return;
    const int len = ipow(nc + 2, dim);
    k_copy_boxes<<<n_blocks((long long)n_slots * len), BLOCK, 0, s>>>(T.stride, slot0, n_slots,
                                                                      len, from, to);
}

void launch_set_tags(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const double *eps)
{
    if (L.n <= 0) return;
    k_set_tags<<<L.n, BLOCK, 0, s>>>(T, L.slot0, ipow(L.nc + 2, dim), eps);
}

int extreme_blocks(int dim, const ListArgs &P)
{
    if (P.n <= 0) return 0;
    const long long th = (long long)P.n * ipow(P.nc, dim);
    const int nb = n_blocks(th);
    return nb < 1024 ? nb : 1024;
}

void launch_extreme_list(int dim, hipStream_t s, const DevTree &T, const ListArgs &P,
                         const double *var, int want_max, double *partial)
{
    const int nb = extreme_blocks(dim, P);
    if (nb == 0) return;
    if (dim == 2)
        k_extreme_list<2><<<nb, BLOCK, 0, s>>>(T, P, var, want_max, partial);
    else
        k_extreme_list<3><<<nb, BLOCK, 0, s>>>(T, P, var, want_max, partial);
}

void launch_reduce_partials(hipStream_t s, const double *partial, int n, int want_max,
                            double *out)
{
    k_reduce_partials<<<1, BLOCK, 0, s>>>(partial, n, want_max, out);
}

void launch_bottom(int dim, hipStream_t s, const DevTree &T, const MgVars &V, const BottomArgs &B)
{
    if (dim == 2)
        k_bottom<2><<<1, BOTTOM_BLOCK, 0, s>>>(T, V, B);
    else
        k_bottom<3><<<1, BOTTOM_BLOCK, 0, s>>>(T, V, B);
}

}  // namespace afg
