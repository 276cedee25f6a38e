// afg_kernels.hip -- gfx950 kernels and launchers of the afivo_amd FAS
// multigrid path.  The per-thread work of every stage is in afg_bodies.h; the
// kernels here map threads to it (one launch per stage and level over the
// HBM pools, or one workgroup for the whole bottom of a V-cycle on LDS).

#include <map>
#include <mutex>

#include "afivo_amd.h"
#include "afg_internal.h"
#include "afg_bodies.h"
#include "afg_lane.h"

namespace afg {

namespace {

constexpr int BLOCK = 256;

inline int n_blocks(long long threads, int block = BLOCK)
{
    return (int)((threads + block - 1) / block);
}

template <typename R>
LevelViewT<R> pool_view(const DevTree &T, const MgVarsT<R> &V)
{
    return LevelViewT<R>{V.phi, V.rhs, V.tmp, V.eps, 0, T.stride, V.lsf, V.bval, V.lsf_mode, V.rstr_mode,
                         V.rhs_rb, V.rb_stride};
}

template <typename R>
LevelViewT<R> var_view(const DevTree &T, R *var)
{
    return LevelViewT<R>{var, nullptr, nullptr, nullptr, 0, T.stride};
}

// ---------------------------------------------------------------------------
// kernels over the HBM pools: one launch per stage and level

template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_fill_gc(DevTree T, LevelArgs L, LevelViewT<R> v)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < fill_threads<DIM>(L.n, NC ? NC : L.nc)) fill_body<DIM, NC>(T, L, v, v, gid);
}

// ghost fill of level L and, in the same launch, tmp = phi - tmp on the
// parents P of the level below (independent work: the fill reads phi only)
template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_fill_diff(DevTree T, LevelArgs L, ListArgs P,
                                                     LevelViewT<R> v)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = fill_threads<DIM>(L.n, NC ? NC : L.nc);
    if (gid < nf) {
        fill_body<DIM, NC>(T, L, v, v, gid);
    } else {
        const int g = gid - nf;
        if (g < P.n * ipow(P.nc + 2, DIM)) diff_body<DIM, 0>(T, P, v, g);
    }
}

template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_diff(DevTree T, ListArgs P, LevelViewT<R> v)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < P.n * ipow((NC ? NC : P.nc) + 2, DIM)) diff_body<DIM, NC>(T, P, v, gid);
}

template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_gsrb_pass(DevTree T, LevelArgs L, LevelViewT<R> v,
                                                     int op_auto, int rb, int fly)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < gsrb_threads<DIM>(L.n, NC ? NC : L.nc))
        gsrb_body<DIM, NC>(T, L, v, v, op_auto, rb, fly, gid);
}

// one colour pass by tiles of boxes in LDS (NC in 2, 4, 8, 16; tile shape V)
template <int DIM, int NC, int FLY, typename R, int V>
__global__ __launch_bounds__((GsrbTile<DIM, NC, V>::THREADS)) void k_gsrb_tile(DevTree T, LevelArgs L,
                                                                              LevelViewT<R> v, int op_auto,
                                                                              int rb)
{
    using G = GsrbTile<DIM, NC, V>;
    __shared__ R lds[G::BOXES * G::LEN];
    R rhs[G::PER];
    const int box0 = xcd_tile(blockIdx.x, gridDim.x) * G::BOXES;
    tile_stage_a<DIM, NC, FLY, V>(T, L, v, v, lds, box0, rb, threadIdx.x, rhs);
    __syncthreads();
    tile_stage_b<DIM, NC, V>(T, L, v, lds, box0, op_auto, rb, threadIdx.x, rhs);
}

// one colour pass of a relaxation block on the colour-split copy of phi (3D
// tiles): MODE 0 the first pass (box arrays in, FLY as k_gsrb_tile), 1 a
// middle one, 2 the last (phi_rb in, box arrays out), 3 the last followed by
// the level's ghost fill (split_fill_push)
// MODE 3 launches may carry, after their n_tiles tile workgroups, workgroups
// that form tmp = phi - tmp over the parents P of the level below
// (correct_children's parent part, which nothing reads before this point).
template <int DIM, int NC, int MODE, int FLY, typename R, int V>
__global__ __launch_bounds__((GsrbTile<DIM, NC, V>::THREADS)) void k_gsrb_split(DevTree T, LevelArgs L,
                                                                               LevelViewT<R> v, R *rbp,
                                                                               int op_auto, int rb, ListArgs P,
                                                                               int n_tiles)
{
    using G = GsrbTile<DIM, NC, V>;
    __shared__ R lds[G::BOXES * G::LEN];
    R rhs[G::PER];
    if constexpr (MODE == 3) {
        if ((int)blockIdx.x >= n_tiles) {   // the whole workgroup
            const long long g = (long long)(blockIdx.x - n_tiles) * G::THREADS + threadIdx.x;
            if (g < (long long)P.n * ipow(P.nc + 2, DIM)) diff_body<DIM, 0>(T, P, v, (int)g);
            return;
        }
    }
    const int box0 = xcd_tile(blockIdx.x, n_tiles) * G::BOXES;
    if (MODE == 0)
        tile_stage_a<DIM, NC, FLY, V>(T, L, v, v, lds, box0, rb, threadIdx.x, rhs);
    else
        split_stage_a<DIM, NC, V>(T, L, v, rbp, lds, box0, rb, threadIdx.x, rhs);
    __syncthreads();
    split_stage_b<DIM, NC, V, MODE>(T, L, v, rbp, lds, box0, op_auto, rb, threadIdx.x, rhs);
    if constexpr (MODE == 3) {
        __syncthreads();
        split_fill_push<DIM, NC, V>(T, L, v, lds, box0, threadIdx.x);
    }
}

// two colour passes (rbA, rbA + 1) by tiles, phi read from v and written to out
template <int DIM, int NC, int FLYA, typename R, int V>
__global__ __launch_bounds__((GsrbTile<DIM, NC, V>::THREADS)) void k_gsrb_pair(DevTree T, LevelArgs L,
                                                                              LevelViewT<R> v, LevelViewT<R> vp,
                                                                              R *out, int op_auto, int rbA)
{
    using G = GsrbTile<DIM, NC, V>;
    __shared__ R lds[G::BOXES * G::LEN];
    R rhsA[G::PER], rhsB[G::PER];
    const int box0 = xcd_tile(blockIdx.x, gridDim.x) * G::BOXES;
    tile_stage_a<DIM, NC, FLYA, V>(T, L, v, vp, lds, box0, rbA, threadIdx.x, rhsA);
    pair_load_rhs<DIM, NC, V>(L, v, box0, rbA + 1, threadIdx.x, rhsB);
    __syncthreads();
    pair_phase1<DIM, NC, FLYA, V>(T, L, v, vp, lds, box0, op_auto, rbA, threadIdx.x, rhsA);
    __syncthreads();
    pair_phase2<DIM, NC, V>(T, L, v, vp, lds, box0, rbA, threadIdx.x);
    __syncthreads();
    pair_phase3<DIM, NC, V>(T, L, v, lds, box0, op_auto, rbA, threadIdx.x, rhsB, out);
}

template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_residual(DevTree T, LevelArgs L, LevelViewT<R> v, int op_auto)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow(NC ? NC : L.nc, DIM)) residual_body<DIM, NC>(T, L, v, v, op_auto, gid);
}

// The mixed solver's FP64 residual with its FP32 image: tmp = rhs - A phi as
// k_residual, and (float) of it into the FP32 rhs pool (same slot layout) and
// its colour-split copy (3D; null: none), which the next correction takes as
// its right-hand side.
template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_residual32(DevTree T, LevelArgs L, LevelViewT<R> v, int op_auto,
                                                      float *r32, float *r32_rb, int rb_stride)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow(NC ? NC : L.nc, DIM)) residual32_body<DIM, NC>(T, L, v, op_auto, r32, r32_rb, rb_stride, gid);
}

template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_uc_child(DevTree T, LevelArgs L, LevelViewT<R> v,
                                                    int op_auto, int keep_child_tmp, int fly)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) / 2, DIM))
        uc_child_body<DIM, NC>(T, L, v, v, op_auto, keep_child_tmp, fly, gid);
}

// update_coarse on FP32 values in FP64 arithmetic (MgVarsT::tau64): child
// part per coarse cell, parent part per cell
template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_uc_child64(DevTree T, LevelArgs L, LevelViewT<R> v,
                                                      int op_auto, int keep_child_tmp, int fly)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) / 2, DIM))
        uc_child_body<DIM, NC, R, double>(T, L, v, v, op_auto, keep_child_tmp, fly, gid);
}

template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_uc_parent64(DevTree T, LevelArgs L, LevelViewT<R> v,
                                                       int op_auto, int copy_tmp)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) + 2, DIM))
        uc_parent_body<DIM, NC, R, double>(T, L, v, v, op_auto, copy_tmp, gid);
}

// update_coarse, child part, by tiles in LDS (NC 8 and 16); FLY 2: the
// level's phi in the colour-split copy rbp (ghosts computed from it)
template <int DIM, int NC, int FLY, typename R>
__global__ __launch_bounds__((UcTile<DIM, NC>::THREADS)) void k_uc_child_tile(DevTree T, LevelArgs L,
                                                                             LevelViewT<R> v, int op_auto,
                                                                             int keep_child_tmp, const R *rbp)
{
    using G = UcTile<DIM, NC>;
    __shared__ R lds[G::BOXES * G::LEN];
    R res[G::PER];
    const int box0 = xcd_tile(blockIdx.x, gridDim.x) * G::BOXES;
    if constexpr (FLY == 2)
        uc_tile_a_split<DIM, NC>(T, L, v, rbp, lds, box0, threadIdx.x);
    else
        uc_tile_a<DIM, NC, FLY>(T, L, v, lds, box0, threadIdx.x);
    __syncthreads();
    uc_tile_b<DIM, NC>(T, L, v, lds, box0, op_auto, keep_child_tmp, threadIdx.x, res);
    __syncthreads();
    uc_tile_d<DIM, NC>(L, lds, box0, threadIdx.x, res);
    __syncthreads();
    uc_tile_restrict<DIM, NC>(T, L, v, lds, box0, threadIdx.x, v.tmp);
}

template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_uc_parent(DevTree T, LevelArgs L, LevelViewT<R> v,
                                                     int op_auto, int copy_tmp)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) + 2, DIM))
        uc_parent_body<DIM, NC>(T, L, v, v, op_auto, copy_tmp, gid);
}

// rbp: the level's phi in the colour-split copy (null: the box arrays)
template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_correct(DevTree T, LevelArgs L, LevelViewT<R> v,
                                                   int op_auto, int fly_diff, R *rbp)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) / 2, DIM))
        correct_body<DIM, NC>(T, L, v, v, op_auto, fly_diff, gid, rbp);
}

// rhs by colour from rhs over the boxes of a level
template <int DIM, typename R>
__global__ __launch_bounds__(BLOCK) void k_split_rhs(DevTree T, LevelArgs L, const R *rhs, R *rhs_rb, int rb_stride)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow(L.nc, DIM)) split_rhs_body<DIM>(T, L, rhs, rhs_rb, rb_stride, gid);
}

// a$D_restrict_box and a$D_prolong2 over the boxes of a level
template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_restrict(DevTree T, LevelArgs L, const R *from, R *to,
                                                    const R *lsf, int rstr_mode)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) / 2, DIM)) restrict_body<DIM, NC>(T, L, from, to, gid, lsf, rstr_mode);
}

template <int DIM, int NC, typename R>
__global__ __launch_bounds__(BLOCK) void k_prolong2(DevTree T, LevelArgs L, const R *from, R *to,
                                                    int add)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < L.n * ipow((NC ? NC : L.nc) / 2, DIM))
        prolong2_body<DIM, NC>(T, L, from, to, add, gid);
}

// tree update: box arrays gathered into a new pool, and the prolongation of
// the boxes the update added
__global__ __launch_bounds__(BLOCK) void k_gather(int stride, int n_slots, const int32_t *src,
                                                  const double *from, double *to)
{
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < (long long)n_slots * (stride >> 1)) gather_body(stride, src, from, to, gid);
}

template <int DIM, int ORDER>
__global__ __launch_bounds__(BLOCK) void k_prolong_list(DevTree T, int nc, int pnc, const int32_t *slots,
                                                        int n, double *var)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < n * ipow(nc / 2, DIM)) prolong_list_body<DIM, ORDER>(T, nc, pnc, slots, var, gid);
}

template <int DIM, int ORDER>
__global__ __launch_bounds__(BLOCK) void k_prolong_level(DevTree T, LevelArgs L, const double *from, double *to,
                                                         int add)
{
    const int per_box = ipow(L.nc / 2, DIM);
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= L.n * per_box) return;
    const int b = gid / per_box;
    prolong_cell<DIM, ORDER>(T, L.nc, L.pnc, L.slot0 + b, from, to, add, gid - b * per_box);
}

// ghost cells of any variable with a choice of methods, and pool arithmetic
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_gc(DevTree T, LevelArgs L, double *var, int rb, int bc)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < fill_threads<DIM>(L.n, L.nc)) gc_body<DIM>(T, L, var, rb, bc, gid);
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_gc2(DevTree T, LevelArgs L, const double *var, double *out, int gstride)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < fill_threads<DIM>(L.n, L.nc)) gc2_body<DIM>(T, L, var, out, gstride, gid);
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_flux(const FluxRec *recs, int n, int nc, int fstride, double *fc,
                                                const double *cyl_rmin, double dr)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < n * (DIM == 2 ? nc : nc * nc)) flux_body<DIM>(recs, nc, fstride, fc, gid, cyl_rmin, dr);
}

__global__ __launch_bounds__(BLOCK) void k_arith(long long n, int op, double a, double b, const double *x,
                                                 double *y)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) arith_body(op, a, b, x, y, i);
}

// whole-array fill and copy over a contiguous slot range
// (a$D_box_clear_cc, m_aX_utils.f90:262-270; a$D_box_copy_cc, :331-339)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_fill_value(int stride, int slot0, int n, int len, R *var,
                                                      R value)
{
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)n * len) return;
    const int b = (int)(gid / len);
    const int q = (int)(gid - (long long)b * len);
    var[(size_t)(slot0 + b) * stride + q] = value;
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_copy_boxes(int stride, int slot0, int n, int len,
                                                      const R *from, R *to)
{
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)n * len) return;
    const int b = (int)(gid / len);
    const int q = (int)(gid - (long long)b * len);
    const size_t o = (size_t)(slot0 + b) * stride + q;
    to[o] = from[o];
}

// FP32 copies of a variable pool and back (mixed-precision solve):
// to[i] = float(from[i]), or to[i] = (add ? to[i] : 0) + double(from[i])
__global__ __launch_bounds__(BLOCK) void k_to_fp32(long long n, const double *from, float *to)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) to[i] = (float)from[i];
}

__global__ __launch_bounds__(BLOCK) void k_from_fp32(long long n, const float *from, double *to, int add)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) to[i] = add ? to[i] + (double)from[i] : (double)from[i];
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

// min and max of a box's whole array (ghosts included), on thread 0
__device__ void box_min_max(const double *e, int len, double &lo, double &hi)
{
    lo = e[0];
    hi = e[0];
    for (int q = threadIdx.x; q < len; q += blockDim.x) {
        lo = e[q] < lo ? e[q] : lo;
        hi = e[q] > hi ? e[q] : hi;
    }
    lo = block_extreme(lo, 0);
    __syncthreads();
    hi = block_extreme(hi, 1);
    __syncthreads();
}

// mg$D_set_box_tag (m_aX_mg.f90:667-706): from min and max of the level set
// (when set) and of eps (when set) over the whole array, one block per box,
// only for boxes whose tag is unset.  A box that gets both a level-set and an
// eps tag (where the reference stops, :700-701) is marked TAG_CONFLICT.
__global__ __launch_bounds__(BLOCK) void k_set_tags(DevTree T, int slot0, int len,
                                                    const double *eps, const double *lsf)
{
    const int slot = slot0 + blockIdx.x;
    if (T.tag[slot] != TAG_UNSET) return;
    bool is_lsf = false, is_eps = false, is_deps = false;
    double lo, hi;
    if (lsf) {
        box_min_max(lsf + (size_t)slot * T.stride, len, lo, hi);
        is_lsf = lo * hi < 0;
    }
    if (eps) {
        box_min_max(eps + (size_t)slot * T.stride, len, lo, hi);
        is_deps = hi > lo;
        if (!is_deps) is_eps = lo < 1 || lo > 1;
    }
    if (threadIdx.x == 0) {
        int8_t tag = TAG_NORMAL;
        if (is_lsf) tag = TAG_LSF;
        if (is_eps) tag = TAG_CEPS;
        if (is_deps) tag = TAG_VEPS;
        if ((int)is_lsf + (int)is_eps + (int)is_deps > 1) tag = TAG_CONFLICT;
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

// max |var| over the leaf interiors of each tree of a forest (a$D_tree_max_cc
// and a$D_tree_min_cc of one tree, max(|min|, |max|), m_aX_utils.f90:468-571).
// Workgroup g takes leaves [chunk[g], chunk[g+1]) of the grouped leaf list
// (all of one tree, nc cells per dimension) and writes its maximum to
// partial[g]; then one workgroup per tree reduces the partials of its chunks
// [tree_chunk[t], tree_chunk[t+1]).
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_tree_maxabs(int stride, int nc, const int32_t *leaves,
                                                       const int32_t *chunk, const double *var,
                                                       double *partial)
{
    const int g = blockIdx.x;
    const int per_box = ipow(nc, DIM);
    const int first = chunk[g];
    const int total = (chunk[g + 1] - first) * per_box;
    double v = 0;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int b = e / per_box;
        const int c = e - b * per_box;
        const int i = c % nc + 1;
        const int j = (c / nc) % nc + 1;
        const int k = DIM == 3 ? c / (nc * nc) + 1 : 0;
        const int s = nc + 2;
        v = pick(v, fabs(var[(size_t)leaves[first + b] * stride + i + s * (j + s * k)]), 1);
    }
    v = block_extreme(v, 1);
    if (threadIdx.x == 0) partial[g] = v;
}

__global__ __launch_bounds__(BLOCK) void k_tree_reduce(const int32_t *tree_chunk, const double *partial,
                                                       double *out)
{
    const int t = blockIdx.x;
    double v = 0;
    for (int q = tree_chunk[t] + threadIdx.x; q < tree_chunk[t + 1]; q += blockDim.x) v = pick(v, partial[q], 1);
    v = block_extreme(v, 1);
    if (threadIdx.x == 0) out[t] = v;
}

// ---------------------------------------------------------------------------
// bottom of the V-cycle in one workgroup, on LDS copies of the small levels
// (bottom_run), and the residual of those levels (bottom_residual_run)

template <int DIM, int AUTO, typename R, int TAU = 0>
__global__ __launch_bounds__(BOTTOM_BLOCK) void k_bottom(DevTree T, MgVarsT<R> V,
                                                          const BottomLevels *__restrict__ B,
                                                          BottomArgs A)
{
    extern __shared__ double lds[];
    // lanes of one wavefront see each other's LDS and global accesses in
    // program order; the fences only keep the compiler from moving them
    const auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    int n_trace = 0;
    auto stamp = [&] {
        if (A.trace && threadIdx.x == 0 && n_trace < A.trace_cap) A.trace[n_trace++] = __builtin_amdgcn_s_memrealtime();
    };
    stamp();
    bottom_run<DIM, AUTO, R, TAU>(T, V, *B, A, lds, threadIdx.x, blockDim.x, warpSize,
                             [&] {
                                 __syncthreads();
                                 stamp();
                             },
                             wave_sync, [](auto v, int mask) { return lane_xor(v, mask); });
    stamp();
}

template <int DIM, typename R>
__global__ __launch_bounds__(BOTTOM_BLOCK) void k_bottom_residual(DevTree T, MgVarsT<R> V,
                                                                   const BottomLevels *__restrict__ B,
                                                                   BottomArgs A)
{
    bottom_residual_run<DIM, R>(T, V, *B, A, threadIdx.x, blockDim.x);
}

template <int DIM, typename R>
__global__ __launch_bounds__(BLOCK) void k_residual_levels(DevTree T, MgVarsT<R> V,
                                                           const BottomLevels *__restrict__ B, int hl)
{
    residual_levels_run<DIM, R>(T, V, *B, hl, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers: dispatch the run-time cell count to a compiled instance

#define AFG_DISPATCH(dim, nc, KERNEL, GRID, ...)                                   \
    do {                                                                           \
        if ((dim) == 2) {                                                          \
            switch (nc) {                                                          \
            case 2: KERNEL<2, 2, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 4: KERNEL<2, 4, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 8: KERNEL<2, 8, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 16: KERNEL<2, 16, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;     \
            default: KERNEL<2, 0, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;      \
            }                                                                      \
        } else {                                                                   \
            switch (nc) {                                                          \
            case 2: KERNEL<3, 2, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 4: KERNEL<3, 4, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 8: KERNEL<3, 8, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;       \
            case 16: KERNEL<3, 16, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;     \
            default: KERNEL<3, 0, R><<<GRID, BLOCK, 0, s>>>(__VA_ARGS__); break;      \
            }                                                                      \
        }                                                                          \
    } while (0)

template <typename R>
void launch_fill_gc(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, R *var)
{
    if (L.n <= 0) return;
    const long long th = dim == 2 ? fill_threads<2>(L.n, L.nc) : fill_threads<3>(L.n, L.nc);
    AFG_DISPATCH(dim, L.nc, k_fill_gc, n_blocks(th), T, L, var_view(T, var));
}

template <typename R>
void launch_fill_diff(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const ListArgs &P, const MgVarsT<R> &V)
{
    if (L.n <= 0) return;
    const long long th = (dim == 2 ? fill_threads<2>(L.n, L.nc) : fill_threads<3>(L.n, L.nc)) +
                         (long long)P.n * ipow(P.nc + 2, dim);
    AFG_DISPATCH(dim, L.nc, k_fill_diff, n_blocks(th), T, L, P, pool_view(T, V));
}

template <typename R>
void launch_diff(int dim, hipStream_t s, const DevTree &T, const ListArgs &P, const MgVarsT<R> &V)
{
    if (P.n <= 0) return;
    const long long th = (long long)P.n * ipow(P.nc + 2, dim);
    AFG_DISPATCH(dim, P.nc, k_diff, n_blocks(th), T, P, pool_view(T, V));
}

template <int DIM, int NC, typename R, int V>
void gsrb_tiles_v(hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &Vs, int rb, int fly)
{
    using G = GsrbTile<DIM, NC, V>;
    const int grid = (L.n + G::BOXES - 1) / G::BOXES;
    if (fly)
        k_gsrb_tile<DIM, NC, 1, R, V><<<grid, G::THREADS, 0, s>>>(T, L, pool_view(T, Vs), Vs.op_auto, rb);
    else
        k_gsrb_tile<DIM, NC, 0, R, V><<<grid, G::THREADS, 0, s>>>(T, L, pool_view(T, Vs), Vs.op_auto, rb);
}

template <int DIM, int NC, typename R>
void gsrb_tiles(hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int rb, int fly)
{
    if constexpr (tile_shapes(NC) > 1) {
        if (L.tile == 1) return gsrb_tiles_v<DIM, NC, R, 1>(s, T, L, V, rb, fly);
        if (L.tile == 2) return gsrb_tiles_v<DIM, NC, R, 2>(s, T, L, V, rb, fly);
    }
    gsrb_tiles_v<DIM, NC, R, 0>(s, T, L, V, rb, fly);
}

template <int DIM, int NC, typename R, int V>
void gsrb_pair_v(hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &Vs, const R *in, R *out,
                 int rbA, int fly)
{
    using G = GsrbTile<DIM, NC, V>;
    const int grid = (L.n + G::BOXES - 1) / G::BOXES;
    const LevelViewT<R> vp = pool_view(T, Vs);
    LevelViewT<R> v = vp;
    v.phi = const_cast<R *>(in);
    if (fly)
        k_gsrb_pair<DIM, NC, 1, R, V><<<grid, G::THREADS, 0, s>>>(T, L, v, vp, out, Vs.op_auto, rbA);
    else
        k_gsrb_pair<DIM, NC, 0, R, V><<<grid, G::THREADS, 0, s>>>(T, L, v, vp, out, Vs.op_auto, rbA);
}

template <int DIM, int NC, typename R>
void gsrb_pair(hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, const R *in, R *out,
               int rbA, int fly)
{
    if (L.tile == 1) return gsrb_pair_v<DIM, NC, R, 1>(s, T, L, V, in, out, rbA, fly);
    if (L.tile == 2) return gsrb_pair_v<DIM, NC, R, 2>(s, T, L, V, in, out, rbA, fly);
    gsrb_pair_v<DIM, NC, R, 0>(s, T, L, V, in, out, rbA, fly);
}

template <typename R>
int launch_gsrb_pair(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V,
                     const R *in, R *out, int rbA, int fly)
{
    if (per_cell_only(T, V)) return 0;
    if (L.n <= 0) return 1;
    switch (dim * 100 + L.nc) {
    case 208: gsrb_pair<2, 8, R>(s, T, L, V, in, out, rbA, fly); return 1;
    case 216: gsrb_pair<2, 16, R>(s, T, L, V, in, out, rbA, fly); return 1;
    case 308: gsrb_pair<3, 8, R>(s, T, L, V, in, out, rbA, fly); return 1;
    case 316: gsrb_pair<3, 16, R>(s, T, L, V, in, out, rbA, fly); return 1;
    default: return 0;
    }
}

template <int NC, typename R, int V>
void gsrb_split_v(hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &Vs, int rb, int mode,
                  int fly, const ListArgs *P)
{
    using G = GsrbTile<3, NC, V>;
    const int n_tiles = (L.n + G::BOXES - 1) / G::BOXES;
    ListArgs Pd{};
    int grid = n_tiles;
    if (mode == 3 && P && P->n > 0) {
        Pd = *P;
        grid += (int)(((long long)Pd.n * ipow(Pd.nc + 2, 3) + G::THREADS - 1) / G::THREADS);
    }
    const LevelViewT<R> v = pool_view(T, Vs);
    if (mode == 0 && fly)
        k_gsrb_split<3, NC, 0, 1, R, V><<<grid, G::THREADS, 0, s>>>(T, L, v, Vs.phi_rb, Vs.op_auto, rb, Pd, n_tiles);
    else if (mode == 0)
        k_gsrb_split<3, NC, 0, 0, R, V><<<grid, G::THREADS, 0, s>>>(T, L, v, Vs.phi_rb, Vs.op_auto, rb, Pd, n_tiles);
    else if (mode == 1)
        k_gsrb_split<3, NC, 1, 1, R, V><<<grid, G::THREADS, 0, s>>>(T, L, v, Vs.phi_rb, Vs.op_auto, rb, Pd, n_tiles);
    else if (mode == 2)
        k_gsrb_split<3, NC, 2, 1, R, V><<<grid, G::THREADS, 0, s>>>(T, L, v, Vs.phi_rb, Vs.op_auto, rb, Pd, n_tiles);
    else
        k_gsrb_split<3, NC, 3, 1, R, V><<<grid, G::THREADS, 0, s>>>(T, L, v, Vs.phi_rb, Vs.op_auto, rb, Pd, n_tiles);
}

template <int NC, typename R>
void gsrb_split(hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int rb, int mode, int fly,
                const ListArgs *P)
{
    if (L.tile == 1) return gsrb_split_v<NC, R, 1>(s, T, L, V, rb, mode, fly, P);
    if (L.tile == 2) return gsrb_split_v<NC, R, 2>(s, T, L, V, rb, mode, fly, P);
    gsrb_split_v<NC, R, 0>(s, T, L, V, rb, mode, fly, P);
}

template <typename R>
int launch_gsrb_split(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int rb,
                      int mode, int fly, const ListArgs *P)
{
    if (dim != 3 || !V.phi_rb || !V.rhs_rb || per_cell_only(T, V)) return 0;
    if (L.n <= 0) return 1;
    switch (L.nc) {
    case 8: gsrb_split<8, R>(s, T, L, V, rb, mode, fly, P); return 1;
    case 16: gsrb_split<16, R>(s, T, L, V, rb, mode, fly, P); return 1;
    default: return 0;
    }
}

template <typename R>
void launch_gsrb_pass(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const MgVarsT<R> &V, int rb, int fly)
{
    if (L.n <= 0) return;
    // cylindrical trees take the per-cell pass (the tiles hold Cartesian operators only)
    switch (per_cell_only(T, V) ? 0 : dim * 100 + L.nc) {
    case 202: gsrb_tiles<2, 2, R>(s, T, L, V, rb, fly); return;
    case 204: gsrb_tiles<2, 4, R>(s, T, L, V, rb, fly); return;
    case 208: gsrb_tiles<2, 8, R>(s, T, L, V, rb, fly); return;
    case 216: gsrb_tiles<2, 16, R>(s, T, L, V, rb, fly); return;
    case 302: gsrb_tiles<3, 2, R>(s, T, L, V, rb, fly); return;
    case 304: gsrb_tiles<3, 4, R>(s, T, L, V, rb, fly); return;
    case 308: gsrb_tiles<3, 8, R>(s, T, L, V, rb, fly); return;
    case 316: gsrb_tiles<3, 16, R>(s, T, L, V, rb, fly); return;
    default: break;
    }
    // other cell counts: one thread per cell of the colour
    const long long th = (long long)L.n * ipow(L.nc, dim) / 2;
    if (dim == 2)
        k_gsrb_pass<2, 0, R><<<n_blocks(th), BLOCK, 0, s>>>(T, L, pool_view(T, V), V.op_auto, rb, fly);
    else
        k_gsrb_pass<3, 0, R><<<n_blocks(th), BLOCK, 0, s>>>(T, L, pool_view(T, V), V.op_auto, rb, fly);
}

template <typename R>
void launch_residual(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const MgVarsT<R> &V)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc, dim);
    AFG_DISPATCH(dim, L.nc, k_residual, n_blocks(th), T, L, pool_view(T, V), V.op_auto);
}

void launch_residual32(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<double> &V,
                       float *r32, float *r32_rb, int rb_stride)
{
    using R = double;
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc, dim);
    AFG_DISPATCH(dim, L.nc, k_residual32, n_blocks(th), T, L, pool_view(T, V), V.op_auto, r32, r32_rb, rb_stride);
}

template <typename R>
void launch_uc_child(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const MgVarsT<R> &V, int keep_child_tmp, int fly, int from_rb)
{
    if (L.n <= 0) return;
    if constexpr (sizeof(R) == 4) {
        if (V.tau64) {
            const long long th = (long long)L.n * ipow(L.nc / 2, dim);
            AFG_DISPATCH(dim, L.nc, k_uc_child64, n_blocks(th), T, L, pool_view(T, V), V.op_auto, keep_child_tmp,
                         fly);
            return;
        }
    }
    if ((L.nc == 8 || L.nc == 16) && !per_cell_only(T, V)) {
        auto go = [&](auto dv, auto nv) {
            constexpr int D = decltype(dv)::value, N = decltype(nv)::value;
            using G = UcTile<D, N>;
            const int grid = (L.n + G::BOXES - 1) / G::BOXES;
            if constexpr (D == 3) {
                if (from_rb) {
                    k_uc_child_tile<D, N, 2, R><<<grid, G::THREADS, 0, s>>>(T, L, pool_view(T, V), V.op_auto,
                                                                            keep_child_tmp, V.phi_rb);
                    return;
                }
            }
            if (fly)
                k_uc_child_tile<D, N, 1, R><<<grid, G::THREADS, 0, s>>>(T, L, pool_view(T, V), V.op_auto,
                                                                        keep_child_tmp, nullptr);
            else
                k_uc_child_tile<D, N, 0, R><<<grid, G::THREADS, 0, s>>>(T, L, pool_view(T, V), V.op_auto,
                                                                        keep_child_tmp, nullptr);
        };
        if (dim == 2)
            L.nc == 8 ? go(IC<2>(), IC<8>()) : go(IC<2>(), IC<16>());
        else
            L.nc == 8 ? go(IC<3>(), IC<8>()) : go(IC<3>(), IC<16>());
        return;
    }
    const long long th = (long long)L.n * ipow(L.nc / 2, dim);
    AFG_DISPATCH(dim, L.nc, k_uc_child, n_blocks(th), T, L, pool_view(T, V), V.op_auto,
                 keep_child_tmp, fly);
}

template <typename R>
void launch_uc_parent(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const MgVarsT<R> &V, int copy_tmp)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc + 2, dim);
    if constexpr (sizeof(R) == 4) {
        if (V.tau64) {
            AFG_DISPATCH(dim, L.nc, k_uc_parent64, n_blocks(th), T, L, pool_view(T, V), V.op_auto, copy_tmp);
            return;
        }
    }
    AFG_DISPATCH(dim, L.nc, k_uc_parent, n_blocks(th), T, L, pool_view(T, V), V.op_auto, copy_tmp);
}

template <typename R>
void launch_correct(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                    const MgVarsT<R> &V, int fly_diff, int to_rb)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc / 2, dim);
    R *rbp = to_rb ? V.phi_rb : nullptr;
    AFG_DISPATCH(dim, L.nc, k_correct, n_blocks(th), T, L, pool_view(T, V), V.op_auto, fly_diff, rbp);
}

template <typename R>
void launch_split_rhs(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const R *rhs, R *rhs_rb,
                      int rb_stride)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc, dim);
    if (dim == 2)
        k_split_rhs<2, R><<<n_blocks(th), BLOCK, 0, s>>>(T, L, rhs, rhs_rb, rb_stride);
    else
        k_split_rhs<3, R><<<n_blocks(th), BLOCK, 0, s>>>(T, L, rhs, rhs_rb, rb_stride);
}

template <typename R>
void launch_restrict(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const R *from, R *to, const R *lsf, int rstr_mode)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc / 2, dim);
    AFG_DISPATCH(dim, L.nc, k_restrict, n_blocks(th), T, L, from, to, lsf, rstr_mode);
}

template <typename R>
void launch_prolong2(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const R *from, R *to, int add)
{
    if (L.n <= 0) return;
    const long long th = (long long)L.n * ipow(L.nc / 2, dim);
    AFG_DISPATCH(dim, L.nc, k_prolong2, n_blocks(th), T, L, from, to, add);
}

template <typename R>
void launch_fill_value(int dim, hipStream_t s, const DevTree &T, int slot0, int n_slots, int nc,
                       R *var, double value)
{
    if (n_slots <= 0) return;
    const int len = ipow(nc + 2, dim);
    k_fill_value<R><<<n_blocks((long long)n_slots * len), BLOCK, 0, s>>>(T.stride, slot0, n_slots,
                                                                         len, var, R(value));
}

template <typename R>
void launch_copy_boxes(int dim, hipStream_t s, const DevTree &T, int slot0, int n_slots, int nc,
                       const R *from, R *to)
{
    if (n_slots <= 0) return;
    const int len = ipow(nc + 2, dim);
    k_copy_boxes<R><<<n_blocks((long long)n_slots * len), BLOCK, 0, s>>>(T.stride, slot0, n_slots,
                                                                         len, from, to);
}

void launch_gather(hipStream_t s, int stride, int n_slots, const int32_t *src, const double *from,
                   double *to)
{
    const long long th = (long long)n_slots * (stride >> 1);
    if (th > 0) k_gather<<<n_blocks(th), BLOCK, 0, s>>>(stride, n_slots, src, from, to);
}

void launch_prolong_list(int dim, hipStream_t s, const DevTree &T, int nc, int pnc, const int32_t *slots,
                         int n, int order, double *var)
{
    if (n <= 0) return;
    const int g = n_blocks((long long)n * ipow(nc / 2, dim));
    if (dim == 2) {
        if (order == 0) k_prolong_list<2, 0><<<g, BLOCK, 0, s>>>(T, nc, pnc, slots, n, var);
        else if (order == 1) k_prolong_list<2, 1><<<g, BLOCK, 0, s>>>(T, nc, pnc, slots, n, var);
        else k_prolong_list<2, 2><<<g, BLOCK, 0, s>>>(T, nc, pnc, slots, n, var);
    } else {
        if (order == 0) k_prolong_list<3, 0><<<g, BLOCK, 0, s>>>(T, nc, pnc, slots, n, var);
        else if (order == 1) k_prolong_list<3, 1><<<g, BLOCK, 0, s>>>(T, nc, pnc, slots, n, var);
        else k_prolong_list<3, 2><<<g, BLOCK, 0, s>>>(T, nc, pnc, slots, n, var);
    }
}

void launch_prolong_level(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const double *from,
                          double *to, int order, int add)
{
    if (L.n <= 0) return;
    const int g = n_blocks((long long)L.n * ipow(L.nc / 2, dim));
    if (dim == 2) {
        if (order == 0) k_prolong_level<2, 0><<<g, BLOCK, 0, s>>>(T, L, from, to, add);
        else if (order == 1) k_prolong_level<2, 1><<<g, BLOCK, 0, s>>>(T, L, from, to, add);
        else k_prolong_level<2, 2><<<g, BLOCK, 0, s>>>(T, L, from, to, add);
    } else {
        if (order == 0) k_prolong_level<3, 0><<<g, BLOCK, 0, s>>>(T, L, from, to, add);
        else if (order == 1) k_prolong_level<3, 1><<<g, BLOCK, 0, s>>>(T, L, from, to, add);
        else k_prolong_level<3, 2><<<g, BLOCK, 0, s>>>(T, L, from, to, add);
    }
}

void launch_gc(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, double *var, int rb, int bc)
{
    if (L.n <= 0) return;
    if (dim == 2) k_gc<2><<<n_blocks(fill_threads<2>(L.n, L.nc)), BLOCK, 0, s>>>(T, L, var, rb, bc);
    else k_gc<3><<<n_blocks(fill_threads<3>(L.n, L.nc)), BLOCK, 0, s>>>(T, L, var, rb, bc);
}

void launch_gc2(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const double *var, double *out,
                int gstride)
{
    if (L.n <= 0) return;
    if (dim == 2) k_gc2<2><<<n_blocks(fill_threads<2>(L.n, L.nc)), BLOCK, 0, s>>>(T, L, var, out, gstride);
    else k_gc2<3><<<n_blocks(fill_threads<3>(L.n, L.nc)), BLOCK, 0, s>>>(T, L, var, out, gstride);
}

void launch_flux(int dim, hipStream_t s, const FluxRec *recs, int n, int nc, int fstride, double *fc,
                 const double *cyl_rmin, double dr)
{
    if (n <= 0) return;
    const int g = n_blocks((long long)n * (dim == 2 ? nc : nc * nc));
    if (dim == 2) k_flux<2><<<g, BLOCK, 0, s>>>(recs, n, nc, fstride, fc, cyl_rmin, dr);
    else k_flux<3><<<g, BLOCK, 0, s>>>(recs, n, nc, fstride, fc, nullptr, dr);
}

void launch_arith(hipStream_t s, long long n, int op, double a, double b, const double *x, double *y)
{
    if (n > 0) k_arith<<<n_blocks(n), BLOCK, 0, s>>>(n, op, a, b, x, y);
}

void launch_to_fp32(hipStream_t s, long long n, const double *from, float *to)
{
    if (n > 0) k_to_fp32<<<n_blocks(n), BLOCK, 0, s>>>(n, from, to);
}

void launch_from_fp32(hipStream_t s, long long n, const float *from, double *to, int add)
{
    if (n > 0) k_from_fp32<<<n_blocks(n), BLOCK, 0, s>>>(n, from, to, add);
}

void launch_set_tags(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const double *eps, const double *lsf)
{
    if (L.n <= 0) return;
    k_set_tags<<<L.n, BLOCK, 0, s>>>(T, L.slot0, ipow(L.nc + 2, dim), eps, lsf);
}

// Scratch memory of the allocation-free reduction (afg_tree_extreme_noalloc):
// the per-block partial results.  One copy per device, held until the
// process ends.  Host code never uses this name as an address: it is a
// host-side stand-in there; scratch_address gives the device address.
static __device__ __align__(256) double scratch[AFG_NOALLOC_MAX_PARTIALS];

hipError_t scratch_address(int device, double **ptr)
{
    static std::mutex mu;
    static std::map<int, double *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *ptr = it->second;
        return hipSuccess;
    }
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    double *p = nullptr;
    e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
    *ptr = p;
    return hipSuccess;
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

void launch_tree_maxabs(int dim, hipStream_t s, int stride, int nc, const int32_t *leaves,
                        const int32_t *chunk, int n_chunks, const int32_t *tree_chunk, int n_trees,
                        const double *var, double *partial, double *out)
{
    if (n_trees <= 0) return;
    if (n_chunks > 0) {
        if (dim == 2)
            k_tree_maxabs<2><<<n_chunks, BLOCK, 0, s>>>(stride, nc, leaves, chunk, var, partial);
        else
            k_tree_maxabs<3><<<n_chunks, BLOCK, 0, s>>>(stride, nc, leaves, chunk, var, partial);
    }
    k_tree_reduce<<<n_trees, BLOCK, 0, s>>>(tree_chunk, partial, out);
}

void launch_reduce_partials(hipStream_t s, const double *partial, int n, int want_max,
                            double *out)
{
    k_reduce_partials<<<1, BLOCK, 0, s>>>(partial, n, want_max, out);
}

// The LDS limit is an attribute of the kernel, shared by every context in
// the process (contexts may be set up from several threads): it only grows,
// to the largest amount any context asked for.
int bottom_prepare(int dim, size_t lds_bytes)
{
    static std::mutex mu;
    static size_t granted[2] = {0, 0};
    std::lock_guard<std::mutex> lock(mu);
    size_t &g = granted[dim == 2 ? 0 : 1];
    if (lds_bytes <= g) return 0;
    const void *fn[6] = {dim == 2 ? (const void *)k_bottom<2, 0, double> : (const void *)k_bottom<3, 0, double>,
                         dim == 2 ? (const void *)k_bottom<2, 1, double> : (const void *)k_bottom<3, 1, double>,
                         dim == 2 ? (const void *)k_bottom<2, 0, float> : (const void *)k_bottom<3, 0, float>,
                         dim == 2 ? (const void *)k_bottom<2, 1, float> : (const void *)k_bottom<3, 1, float>,
                         dim == 2 ? (const void *)k_bottom<2, 0, float, 1> : (const void *)k_bottom<3, 0, float, 1>,
                         dim == 2 ? (const void *)k_bottom<2, 1, float, 1> : (const void *)k_bottom<3, 1, float, 1>};
    for (const void *f : fn)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
            return -1;
    g = lds_bytes;
    return 0;
}

template <typename R>
void launch_bottom(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                   const BottomLevels *B, const BottomArgs &A, size_t lds_bytes, int threads)
{
    if constexpr (sizeof(R) == 4) {
        if (A.tau64) {
            if (dim == 2) {
                if (V.op_auto)
                    k_bottom<2, 1, R, 1><<<1, threads, lds_bytes, s>>>(T, V, B, A);
                else
                    k_bottom<2, 0, R, 1><<<1, threads, lds_bytes, s>>>(T, V, B, A);
            } else {
                if (V.op_auto)
                    k_bottom<3, 1, R, 1><<<1, threads, lds_bytes, s>>>(T, V, B, A);
                else
                    k_bottom<3, 0, R, 1><<<1, threads, lds_bytes, s>>>(T, V, B, A);
            }
            return;
        }
    }
    if (dim == 2) {
        if (V.op_auto)
            k_bottom<2, 1, R><<<1, threads, lds_bytes, s>>>(T, V, B, A);
        else
            k_bottom<2, 0, R><<<1, threads, lds_bytes, s>>>(T, V, B, A);
    } else {
        if (V.op_auto)
            k_bottom<3, 1, R><<<1, threads, lds_bytes, s>>>(T, V, B, A);
        else
            k_bottom<3, 0, R><<<1, threads, lds_bytes, s>>>(T, V, B, A);
    }
}

template <typename R>
void launch_residual_levels(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                            const BottomLevels *B, int hl, int64_t max_cells)
{
    if (max_cells <= 0) return;
    if (dim == 2)
        k_residual_levels<2, R><<<n_blocks(max_cells), BLOCK, 0, s>>>(T, V, B, hl);
    else
        k_residual_levels<3, R><<<n_blocks(max_cells), BLOCK, 0, s>>>(T, V, B, hl);
}

template <typename R>
void launch_bottom_residual(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                            const BottomLevels *B, const BottomArgs &A, int threads)
{
    if (dim == 2)
        k_bottom_residual<2, R><<<1, threads, 0, s>>>(T, V, B, A);
    else
        k_bottom_residual<3, R><<<1, threads, 0, s>>>(T, V, B, A);
}

// the two value types of the launchers
#define AFG_INSTANTIATE(R)                                                                               \
    template void launch_fill_gc<R>(int, hipStream_t, const DevTree &, const LevelArgs &, R *);         \
    template int launch_gsrb_split<R>(int, hipStream_t, const DevTree &, const LevelArgs &,            \
                                      const MgVarsT<R> &, int, int, int, const ListArgs *);            \
    template int launch_gsrb_pair<R>(int, hipStream_t, const DevTree &, const LevelArgs &,             \
                                     const MgVarsT<R> &, const R *, R *, int, int);                    \
    template void launch_fill_diff<R>(int, hipStream_t, const DevTree &, const LevelArgs &,            \
                                      const ListArgs &, const MgVarsT<R> &);                           \
    template void launch_diff<R>(int, hipStream_t, const DevTree &, const ListArgs &,                  \
                                 const MgVarsT<R> &);                                                  \
    template void launch_gsrb_pass<R>(int, hipStream_t, const DevTree &, const LevelArgs &,            \
                                      const MgVarsT<R> &, int, int);                                   \
    template void launch_residual<R>(int, hipStream_t, const DevTree &, const LevelArgs &,             \
                                     const MgVarsT<R> &);                                              \
    template void launch_uc_child<R>(int, hipStream_t, const DevTree &, const LevelArgs &,             \
                                     const MgVarsT<R> &, int, int, int);                               \
    template void launch_uc_parent<R>(int, hipStream_t, const DevTree &, const LevelArgs &,            \
                                      const MgVarsT<R> &, int);                                        \
    template void launch_correct<R>(int, hipStream_t, const DevTree &, const LevelArgs &,              \
                                    const MgVarsT<R> &, int, int);                                     \
    template void launch_restrict<R>(int, hipStream_t, const DevTree &, const LevelArgs &, const R *, R *, const R *, int); \
    template void launch_split_rhs<R>(int, hipStream_t, const DevTree &, const LevelArgs &, const R *, R *, int); \
    template void launch_prolong2<R>(int, hipStream_t, const DevTree &, const LevelArgs &, const R *, R *, \
                                     int);                                                             \
    template void launch_fill_value<R>(int, hipStream_t, const DevTree &, int, int, int, R *, double);  \
    template void launch_copy_boxes<R>(int, hipStream_t, const DevTree &, int, int, int, const R *, R *); \
    template void launch_bottom<R>(int, hipStream_t, const DevTree &, const MgVarsT<R> &,              \
                                   const BottomLevels *, const BottomArgs &, size_t, int);             \
    template void launch_bottom_residual<R>(int, hipStream_t, const DevTree &, const MgVarsT<R> &,     \
                                            const BottomLevels *, const BottomArgs &, int);            \
    template void launch_residual_levels<R>(int, hipStream_t, const DevTree &, const MgVarsT<R> &,     \
                                            const BottomLevels *, int, int64_t);
AFG_INSTANTIATE(double)
AFG_INSTANTIATE(float)

}  // namespace afg
