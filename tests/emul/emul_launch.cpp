// CPU emulation of the afivo_amd kernel launchers.  TEST INFRASTRUCTURE ONLY.
//
// Each launcher runs the threads of its stage one after the other through
// the same per-thread bodies the HIP kernels use (afivo_amd/csrc/afg_bodies.h,
// compiled here for the host).  A stage is race-free, so any thread order
// gives the kernel's result; the bottom kernel runs as one thread whose
// workgroup barriers are no-ops.  Built with -ffp-contract=off, the host
// arithmetic rounds like the device's, so results compare bit for bit.
#define AFG_DEV inline
#define AFG_HD
#include "afivo_amd.h"
#include "afg_internal.h"
#include "afg_bodies.h"
#include "emul.h"

#include <algorithm>
#include <cmath>
#include <vector>
#include <cfloat>

namespace afg {

namespace {

constexpr int BLOCK = 256;

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

// Run body<DIM, NC>(gid) for gid in [0, n(DIM)) with NC dispatched like the
// device launchers do.
template <typename Body, typename Count>
void run_stage(int dim, int nc, Count count, Body body)
{
    if (dim == 2) {
        nc_dispatch(nc, [&](auto ncv) {
            const long long n = count(std::integral_constant<int, 2>());
            for (long long g = 0; g < n; g++) body(std::integral_constant<int, 2>(), ncv, (int)g);
        });
    } else {
        nc_dispatch(nc, [&](auto ncv) {
            const long long n = count(std::integral_constant<int, 3>());
            for (long long g = 0; g < n; g++) body(std::integral_constant<int, 3>(), ncv, (int)g);
        });
    }
}

#define DIMV decltype(d)::value
#define NCV decltype(c)::value

}  // namespace

template <typename R>
void launch_fill_gc(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, R *var)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        const LevelViewT<R> v = var_view(T, var);
        run_stage(dim, L.nc, [&](auto d) { return fill_threads<DIMV>(L.n, L.nc); },
                  [&](auto d, auto c, int g) { fill_body<DIMV, NCV>(T, L, v, v, g); });
    });
}

template <typename R>
void launch_fill_diff(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const ListArgs &P, const MgVarsT<R> &V)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        const LevelViewT<R> v = pool_view(T, V);
        run_stage(dim, L.nc, [&](auto d) { return fill_threads<DIMV>(L.n, L.nc); },
                  [&](auto d, auto c, int g) { fill_body<DIMV, NCV>(T, L, v, v, g); });
        run_stage(dim, P.nc, [&](auto d) { return (long long)P.n * ipow(P.nc + 2, DIMV); },
                  [&](auto d, auto c, int g) { diff_body<DIMV, 0>(T, P, v, g); });
    });
}

template <typename R>
void launch_diff(int dim, hipStream_t s, const DevTree &T, const ListArgs &P, const MgVarsT<R> &V)
{
    if (P.n <= 0) return;
    emul_submit(s, [=] {
        const LevelViewT<R> v = pool_view(T, V);
        run_stage(dim, P.nc, [&](auto d) { return (long long)P.n * ipow(P.nc + 2, DIMV); },
                  [&](auto d, auto c, int g) { diff_body<DIMV, NCV>(T, P, v, g); });
    });
}

// the tiled pass of the device: workgroups one after the other, each stage
// over all threads of the workgroup before the next
template <int DIM, int NC, typename R, int SH>
void gsrb_tiles_v(const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int rb, int fly)
{
    using G = GsrbTile<DIM, NC, SH>;
    const LevelViewT<R> v = pool_view(T, V);
    const int grid = (L.n + G::BOXES - 1) / G::BOXES;
    R *lds = (R *)emul_alloc(sizeof(R) * G::BOXES * G::LEN);
    std::vector<R> rhs((size_t)G::THREADS * G::PER);
    for (int blk = 0; blk < grid; blk++) {
        const int box0 = xcd_tile(blk, grid) * G::BOXES;
        for (int t = 0; t < G::THREADS; t++) {
            if (fly)
                tile_stage_a<DIM, NC, 1, SH>(T, L, v, v, lds, box0, rb, t, &rhs[(size_t)t * G::PER]);
            else
                tile_stage_a<DIM, NC, 0, SH>(T, L, v, v, lds, box0, rb, t, &rhs[(size_t)t * G::PER]);
        }
        for (int t = 0; t < G::THREADS; t++)
            tile_stage_b<DIM, NC, SH>(T, L, v, lds, box0, V.op_auto, rb, t, &rhs[(size_t)t * G::PER]);
    }
    emul_free(lds);
}

template <int DIM, int NC, typename R>
void gsrb_tiles(const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int rb, int fly)
{
    if constexpr (tile_shapes(NC) > 1) {
        if (L.tile == 1) return gsrb_tiles_v<DIM, NC, R, 1>(T, L, V, rb, fly);
        if (L.tile == 2) return gsrb_tiles_v<DIM, NC, R, 2>(T, L, V, rb, fly);
    }
    gsrb_tiles_v<DIM, NC, R, 0>(T, L, V, rb, fly);
}

// the colour passes on the colour-split copy of phi: workgroups one after
// the other, each stage over all threads of the workgroup before the next
template <int NC, typename R, int SH>
void gsrb_split_v(const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int rb, int mode, int fly,
                  const ListArgs *P)
{
    using G = GsrbTile<3, NC, SH>;
    const LevelViewT<R> v = pool_view(T, V);
    const int grid = (L.n + G::BOXES - 1) / G::BOXES;
    R *lds = (R *)emul_alloc(sizeof(R) * G::BOXES * G::LEN);
    std::vector<R> rhs((size_t)G::THREADS * G::PER);
    for (int blk = 0; blk < grid; blk++) {
        const int box0 = xcd_tile(blk, grid) * G::BOXES;
        for (int t = 0; t < G::THREADS; t++) {
            R *r = &rhs[(size_t)t * G::PER];
            if (mode == 0 && fly)
                tile_stage_a<3, NC, 1, SH>(T, L, v, v, lds, box0, rb, t, r);
            else if (mode == 0)
                tile_stage_a<3, NC, 0, SH>(T, L, v, v, lds, box0, rb, t, r);
            else
                split_stage_a<3, NC, SH>(T, L, v, V.phi_rb, lds, box0, rb, t, r);
        }
        for (int t = 0; t < G::THREADS; t++) {
            const R *r = &rhs[(size_t)t * G::PER];
            if (mode == 0)
                split_stage_b<3, NC, SH, 0>(T, L, v, V.phi_rb, lds, box0, V.op_auto, rb, t, r);
            else if (mode == 1)
                split_stage_b<3, NC, SH, 1>(T, L, v, V.phi_rb, lds, box0, V.op_auto, rb, t, r);
            else if (mode == 2)
                split_stage_b<3, NC, SH, 2>(T, L, v, V.phi_rb, lds, box0, V.op_auto, rb, t, r);
            else
                split_stage_b<3, NC, SH, 3>(T, L, v, V.phi_rb, lds, box0, V.op_auto, rb, t, r);
        }
        if (mode == 3)
            for (int t = 0; t < G::THREADS; t++) split_fill_push<3, NC, SH>(T, L, v, lds, box0, t);
    }
    if (mode == 3 && P)
        for (long long g = 0; g < (long long)P->n * ipow(P->nc + 2, 3); g++) diff_body<3, 0>(T, *P, v, (int)g);
    emul_free(lds);
}

template <typename R>
int launch_gsrb_split(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int rb,
                      int mode, int fly, const ListArgs *P)
{
    if (dim != 3 || !V.phi_rb || !V.rhs_rb || per_cell_only(T, V)) return 0;
    if (L.n <= 0) return 1;
    if (L.nc != 8 && L.nc != 16) return 0;
    const bool has_p = P != nullptr;
    const ListArgs Pc = has_p ? *P : ListArgs{};
    emul_submit(s, [=] {
        const ListArgs *pp = has_p ? &Pc : nullptr;
        auto go = [&](auto ncv) {
            constexpr int NC = decltype(ncv)::value;
            if (L.tile == 1)
                gsrb_split_v<NC, R, 1>(T, L, V, rb, mode, fly, pp);
            else if (L.tile == 2)
                gsrb_split_v<NC, R, 2>(T, L, V, rb, mode, fly, pp);
            else
                gsrb_split_v<NC, R, 0>(T, L, V, rb, mode, fly, pp);
        };
        if (L.nc == 8)
            go(IC<8>());
        else
            go(IC<16>());
    });
    return 1;
}

// the tiled child stage of the device: workgroups one after the other, each
// phase over all threads of the workgroup before the next
template <int DIM, int NC, typename R>
void uc_tiles(const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int keep_child_tmp, int fly, int from_rb)
{
    using G = UcTile<DIM, NC>;
    const LevelViewT<R> v = pool_view(T, V);
    const int grid = (L.n + G::BOXES - 1) / G::BOXES;
    R *lds = (R *)emul_alloc(sizeof(R) * G::BOXES * G::LEN);
    std::vector<R> res((size_t)G::THREADS * G::PER);
    for (int blk = 0; blk < grid; blk++) {
        const int box0 = xcd_tile(blk, grid) * G::BOXES;
        for (int t = 0; t < G::THREADS; t++) {
            if (DIM == 3 && from_rb)
                uc_tile_a_split<DIM, NC>(T, L, v, V.phi_rb, lds, box0, t);
            else if (fly)
                uc_tile_a<DIM, NC, 1>(T, L, v, lds, box0, t);
            else
                uc_tile_a<DIM, NC, 0>(T, L, v, lds, box0, t);
        }
        for (int t = 0; t < G::THREADS; t++)
            uc_tile_b<DIM, NC>(T, L, v, lds, box0, V.op_auto, keep_child_tmp, t, &res[(size_t)t * G::PER]);
        for (int t = 0; t < G::THREADS; t++) uc_tile_d<DIM, NC>(L, lds, box0, t, &res[(size_t)t * G::PER]);
        for (int t = 0; t < G::THREADS; t++) uc_tile_restrict<DIM, NC>(T, L, v, lds, box0, t, v.tmp);
    }
    emul_free(lds);
}

// the fused pass pair of the device: workgroups one after the other, each
// phase over all threads of the workgroup before the next
template <int DIM, int NC, typename R, int SH>
void gsrb_pair_v(const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, const R *in, R *out, int rbA, int fly)
{
    using G = GsrbTile<DIM, NC, SH>;
    const LevelViewT<R> vp = pool_view(T, V);
    LevelViewT<R> v = vp;
    v.phi = const_cast<R *>(in);
    const int grid = (L.n + G::BOXES - 1) / G::BOXES;
    R *lds = (R *)emul_alloc(sizeof(R) * G::BOXES * G::LEN);
    std::vector<R> rA((size_t)G::THREADS * G::PER), rB((size_t)G::THREADS * G::PER);
    for (int blk = 0; blk < grid; blk++) {
        const int box0 = xcd_tile(blk, grid) * G::BOXES;
        for (int t = 0; t < G::THREADS; t++) {
            if (fly)
                tile_stage_a<DIM, NC, 1, SH>(T, L, v, vp, lds, box0, rbA, t, &rA[(size_t)t * G::PER]);
            else
                tile_stage_a<DIM, NC, 0, SH>(T, L, v, vp, lds, box0, rbA, t, &rA[(size_t)t * G::PER]);
            pair_load_rhs<DIM, NC, SH>(L, v, box0, rbA + 1, t, &rB[(size_t)t * G::PER]);
        }
        for (int t = 0; t < G::THREADS; t++) {
            if (fly)
                pair_phase1<DIM, NC, 1, SH>(T, L, v, vp, lds, box0, V.op_auto, rbA, t, &rA[(size_t)t * G::PER]);
            else
                pair_phase1<DIM, NC, 0, SH>(T, L, v, vp, lds, box0, V.op_auto, rbA, t, &rA[(size_t)t * G::PER]);
        }
        for (int t = 0; t < G::THREADS; t++) pair_phase2<DIM, NC, SH>(T, L, v, vp, lds, box0, rbA, t);
        for (int t = 0; t < G::THREADS; t++)
            pair_phase3<DIM, NC, SH>(T, L, v, lds, box0, V.op_auto, rbA, t, &rB[(size_t)t * G::PER], out);
    }
    emul_free(lds);
}

template <int DIM, int NC, typename R>
void gsrb_pair(const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, const R *in, R *out, int rbA, int fly)
{
    if (L.tile == 1) return gsrb_pair_v<DIM, NC, R, 1>(T, L, V, in, out, rbA, fly);
    if (L.tile == 2) return gsrb_pair_v<DIM, NC, R, 2>(T, L, V, in, out, rbA, fly);
    gsrb_pair_v<DIM, NC, R, 0>(T, L, V, in, out, rbA, fly);
}

template <typename R>
int launch_gsrb_pair(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V,
                     const R *in, R *out, int rbA, int fly)
{
    if (per_cell_only(T, V)) return 0;
    if (L.n <= 0) return 1;
    const int key = dim * 100 + L.nc;
    if (key != 208 && key != 216 && key != 308 && key != 316) return 0;
    emul_submit(s, [=] {
        switch (key) {
        case 208: gsrb_pair<2, 8, R>(T, L, V, in, out, rbA, fly); break;
        case 216: gsrb_pair<2, 16, R>(T, L, V, in, out, rbA, fly); break;
        case 308: gsrb_pair<3, 8, R>(T, L, V, in, out, rbA, fly); break;
        default: gsrb_pair<3, 16, R>(T, L, V, in, out, rbA, fly); break;
        }
    });
    return 1;
}

template <typename R>
void launch_gsrb_pass(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const MgVarsT<R> &V, int rb, int fly)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        switch (per_cell_only(T, V) ? 0 : dim * 100 + L.nc) {
        case 202: gsrb_tiles<2, 2, R>(T, L, V, rb, fly); return;
        case 204: gsrb_tiles<2, 4, R>(T, L, V, rb, fly); return;
        case 208: gsrb_tiles<2, 8, R>(T, L, V, rb, fly); return;
        case 216: gsrb_tiles<2, 16, R>(T, L, V, rb, fly); return;
        case 302: gsrb_tiles<3, 2, R>(T, L, V, rb, fly); return;
        case 304: gsrb_tiles<3, 4, R>(T, L, V, rb, fly); return;
        case 308: gsrb_tiles<3, 8, R>(T, L, V, rb, fly); return;
        case 316: gsrb_tiles<3, 16, R>(T, L, V, rb, fly); return;
        default: break;
        }
        const LevelViewT<R> v = pool_view(T, V);
        run_stage(dim, 0, [&](auto d) { return gsrb_threads<DIMV>(L.n, L.nc); },
                  [&](auto d, auto c, int g) { gsrb_body<DIMV, 0>(T, L, v, v, V.op_auto, rb, fly, g); });
    });
}

template <typename R>
void launch_residual(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        const LevelViewT<R> v = pool_view(T, V);
        run_stage(dim, L.nc, [&](auto d) { return (long long)L.n * ipow(L.nc, DIMV); },
                  [&](auto d, auto c, int g) { residual_body<DIMV, NCV>(T, L, v, v, V.op_auto, g); });
    });
}

void launch_residual32(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<double> &V,
                       float *r32, float *r32_rb, int rb_stride)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        const LevelViewT<double> v = pool_view(T, V);
        run_stage(dim, L.nc, [&](auto d) { return (long long)L.n * ipow(L.nc, DIMV); },
                  [&](auto d, auto c, int g) { residual32_body<DIMV, NCV>(T, L, v, V.op_auto, r32, r32_rb, rb_stride, g); });
    });
}

template <typename R>
void launch_uc_child(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const MgVarsT<R> &V, int keep_child_tmp, int fly, int from_rb)
{
    if (L.n <= 0) return;
    if constexpr (sizeof(R) == 4) {
        if (V.tau64) {
            emul_submit(s, [=] {
                const LevelViewT<R> v = pool_view(T, V);
                run_stage(dim, L.nc, [&](auto d) { return (long long)L.n * ipow(L.nc / 2, DIMV); },
                          [&](auto d, auto c, int g) {
                              uc_child_body<DIMV, NCV, R, double>(T, L, v, v, V.op_auto, keep_child_tmp, fly, g);
                          });
            });
            return;
        }
    }
    if ((L.nc == 8 || L.nc == 16) && !per_cell_only(T, V)) {
        emul_submit(s, [=] {
            if (dim == 2)
                L.nc == 8 ? uc_tiles<2, 8, R>(T, L, V, keep_child_tmp, fly, 0) : uc_tiles<2, 16, R>(T, L, V, keep_child_tmp, fly, 0);
            else
                L.nc == 8 ? uc_tiles<3, 8, R>(T, L, V, keep_child_tmp, fly, from_rb)
                          : uc_tiles<3, 16, R>(T, L, V, keep_child_tmp, fly, from_rb);
        });
        return;
    }
    emul_submit(s, [=] {
        const LevelViewT<R> v = pool_view(T, V);
        run_stage(dim, L.nc, [&](auto d) { return (long long)L.n * ipow(L.nc / 2, DIMV); },
                  [&](auto d, auto c, int g) {
                      uc_child_body<DIMV, NCV>(T, L, v, v, V.op_auto, keep_child_tmp, fly, g);
                  });
    });
}

template <typename R>
void launch_uc_parent(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const MgVarsT<R> &V, int copy_tmp)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        const LevelViewT<R> v = pool_view(T, V);
        run_stage(dim, L.nc, [&](auto d) { return (long long)L.n * ipow(L.nc + 2, DIMV); },
                  [&](auto d, auto c, int g) {
                      if constexpr (sizeof(R) == 4) {
                          if (V.tau64) {
                              uc_parent_body<DIMV, NCV, R, double>(T, L, v, v, V.op_auto, copy_tmp, g);
                              return;
                          }
                      }
                      uc_parent_body<DIMV, NCV>(T, L, v, v, V.op_auto, copy_tmp, g);
                  });
    });
}

template <typename R>
void launch_correct(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                    const MgVarsT<R> &V, int fly_diff, int to_rb)
{
    if (L.n <= 0) return;
    R *rbp = to_rb ? V.phi_rb : nullptr;
    emul_submit(s, [=] {
        const LevelViewT<R> v = pool_view(T, V);
        run_stage(dim, L.nc, [&](auto d) { return (long long)L.n * ipow(L.nc / 2, DIMV); },
                  [&](auto d, auto c, int g) { correct_body<DIMV, NCV>(T, L, v, v, V.op_auto, fly_diff, g, rbp); });
    });
}

template <typename R>
void launch_split_rhs(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const R *rhs, R *rhs_rb,
                      int rb_stride)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        const long long n = (long long)L.n * ipow(L.nc, dim);
        for (long long g = 0; g < n; g++) {
            if (dim == 2)
                split_rhs_body<2>(T, L, rhs, rhs_rb, rb_stride, (int)g);
            else
                split_rhs_body<3>(T, L, rhs, rhs_rb, rb_stride, (int)g);
        }
    });
}

template <typename R>
void launch_restrict(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const R *from, R *to, const R *lsf, int rstr_mode)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        run_stage(dim, L.nc, [&](auto d) { return (long long)L.n * ipow(L.nc / 2, DIMV); },
                  [&](auto d, auto c, int g) { restrict_body<DIMV, NCV>(T, L, from, to, g, lsf, rstr_mode); });
    });
}

template <typename R>
void launch_prolong2(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const R *from, R *to, int add)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        run_stage(dim, L.nc, [&](auto d) { return (long long)L.n * ipow(L.nc / 2, DIMV); },
                  [&](auto d, auto c, int g) { prolong2_body<DIMV, NCV>(T, L, from, to, add, g); });
    });
}

template <typename R>
void launch_fill_value(int dim, hipStream_t s, const DevTree &T, int slot0, int n_slots, int nc,
                       R *var, double value)
{
    if (n_slots <= 0) return;
    emul_submit(s, [=] {
        const int len = ipow(nc + 2, dim);
        for (int b = 0; b < n_slots; b++)
            for (int q = 0; q < len; q++) var[(size_t)(slot0 + b) * T.stride + q] = R(value);
    });
}

template <typename R>
void launch_copy_boxes(int dim, hipStream_t s, const DevTree &T, int slot0, int n_slots, int nc,
                       const R *from, R *to)
{
    if (n_slots <= 0) return;
    emul_submit(s, [=] {
        const int len = ipow(nc + 2, dim);
        for (int b = 0; b < n_slots; b++)
            for (int q = 0; q < len; q++) {
                const size_t o = (size_t)(slot0 + b) * T.stride + q;
                to[o] = from[o];
            }
    });
}

void launch_gather(hipStream_t s, int stride, int n_slots, const int32_t *src, const double *from,
                   double *to)
{
    const long long th = (long long)n_slots * (stride >> 1);
    if (th > 0)
        emul_submit(s, [=] { for (long long g = 0; g < th; g++) gather_body(stride, src, from, to, g); });
}

void launch_prolong_list(int dim, hipStream_t s, const DevTree &T, int nc, int pnc, const int32_t *slots,
                         int n, int order, double *var)
{
    if (n <= 0) return;
    emul_submit(s, [=] {
        const int th = n * ipow(nc / 2, dim);
        for (int g = 0; g < th; g++) {
            if (dim == 2) {
                if (order == 0) prolong_list_body<2, 0>(T, nc, pnc, slots, var, g);
                else if (order == 1) prolong_list_body<2, 1>(T, nc, pnc, slots, var, g);
                else prolong_list_body<2, 2>(T, nc, pnc, slots, var, g);
            } else {
                if (order == 0) prolong_list_body<3, 0>(T, nc, pnc, slots, var, g);
                else if (order == 1) prolong_list_body<3, 1>(T, nc, pnc, slots, var, g);
                else prolong_list_body<3, 2>(T, nc, pnc, slots, var, g);
            }
        }
    });
}

void launch_prolong_level(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const double *from,
                          double *to, int order, int add)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        const int per_box = ipow(L.nc / 2, dim);
        for (int g = 0; g < L.n * per_box; g++) {
            const int b = g / per_box, c = g - b * per_box, slot = L.slot0 + b;
            if (dim == 2) {
                if (order == 0) prolong_cell<2, 0>(T, L.nc, L.pnc, slot, from, to, add, c);
                else if (order == 1) prolong_cell<2, 1>(T, L.nc, L.pnc, slot, from, to, add, c);
                else prolong_cell<2, 2>(T, L.nc, L.pnc, slot, from, to, add, c);
            } else {
                if (order == 0) prolong_cell<3, 0>(T, L.nc, L.pnc, slot, from, to, add, c);
                else if (order == 1) prolong_cell<3, 1>(T, L.nc, L.pnc, slot, from, to, add, c);
                else prolong_cell<3, 2>(T, L.nc, L.pnc, slot, from, to, add, c);
            }
        }
    });
}

void launch_gc(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, double *var, int rb, int bc)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        if (dim == 2)
            for (int g = 0; g < fill_threads<2>(L.n, L.nc); g++) gc_body<2>(T, L, var, rb, bc, g);
        else
            for (int g = 0; g < fill_threads<3>(L.n, L.nc); g++) gc_body<3>(T, L, var, rb, bc, g);
    });
}

void launch_gc2(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const double *var, double *out,
                int gstride)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        if (dim == 2)
            for (int g = 0; g < fill_threads<2>(L.n, L.nc); g++) gc2_body<2>(T, L, var, out, gstride, g);
        else
            for (int g = 0; g < fill_threads<3>(L.n, L.nc); g++) gc2_body<3>(T, L, var, out, gstride, g);
    });
}

void launch_flux(int dim, hipStream_t s, const FluxRec *recs, int n, int nc, int fstride, double *fc,
                 const double *cyl_rmin, double dr)
{
    if (n <= 0) return;
    emul_submit(s, [=] {
        const int th = n * (dim == 2 ? nc : nc * nc);
        for (int g = 0; g < th; g++) {
            if (dim == 2) flux_body<2>(recs, nc, fstride, fc, g, cyl_rmin, dr);
            else flux_body<3>(recs, nc, fstride, fc, g);
        }
    });
}

void launch_arith(hipStream_t s, long long n, int op, double a, double b, const double *x, double *y)
{
    if (n > 0) emul_submit(s, [=] { for (long long i = 0; i < n; i++) arith_body(op, a, b, x, y, i); });
}

void launch_to_fp32(hipStream_t s, long long n, const double *from, float *to)
{
    if (n > 0) emul_submit(s, [=] { for (long long i = 0; i < n; i++) to[i] = (float)from[i]; });
}

void launch_from_fp32(hipStream_t s, long long n, const float *from, double *to, int add)
{
    if (n > 0)
        emul_submit(s, [=] {
            for (long long i = 0; i < n; i++) to[i] = add ? to[i] + (double)from[i] : (double)from[i];
        });
}

void launch_set_tags(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const double *eps,
                     const double *lsf)
{
    if (L.n <= 0) return;
    emul_submit(s, [=] {
        const int len = ipow(L.nc + 2, dim);
        auto min_max = [&](const double *e, double &lo, double &hi) {
            lo = hi = e[0];
            for (int q = 0; q < len; q++) {
                lo = e[q] < lo ? e[q] : lo;
                hi = e[q] > hi ? e[q] : hi;
            }
        };
        for (int b = 0; b < L.n; b++) {
            const int slot = L.slot0 + b;
            if (T.tag[slot] != TAG_UNSET) continue;
            bool is_lsf = false, is_eps = false, is_deps = false;
            double lo, hi;
            if (lsf) {
                min_max(lsf + (size_t)slot * T.stride, lo, hi);
                is_lsf = lo * hi < 0;
            }
            if (eps) {
                min_max(eps + (size_t)slot * T.stride, lo, hi);
                is_deps = hi > lo;
                if (!is_deps) is_eps = lo < 1 || lo > 1;
            }
            int8_t tag = TAG_NORMAL;
            if (is_lsf) tag = TAG_LSF;
            if (is_eps) tag = TAG_CEPS;
            if (is_deps) tag = TAG_VEPS;
            if ((int)is_lsf + (int)is_eps + (int)is_deps > 1) tag = TAG_CONFLICT;
            T.tag[slot] = tag;
        }
    });
}

int extreme_blocks(int dim, const ListArgs &P)
{
    if (P.n <= 0) return 0;
    const long long th = (long long)P.n * ipow(P.nc, dim);
    const int nb = (int)((th + BLOCK - 1) / BLOCK);
    return nb < 1024 ? nb : 1024;
}

void launch_extreme_list(int dim, hipStream_t s, const DevTree &T, const ListArgs &P,
                         const double *var, int want_max, double *partial)
{
    const int nb = extreme_blocks(dim, P);
    if (nb == 0) return;
    emul_submit(s, [=] {
        double v = want_max ? -DBL_MAX : DBL_MAX;
        const int nc = P.nc, sp = nc + 2;
        for (int b = 0; b < P.n; b++)
            for (int k = dim == 3 ? 1 : 0; k <= (dim == 3 ? nc : 0); k++)
                for (int j = 1; j <= nc; j++)
                    for (int i = 1; i <= nc; i++) {
                        const double x = var[(size_t)P.slots[b] * T.stride + i + sp * (j + sp * k)];
                        v = want_max ? (x > v ? x : v) : (x < v ? x : v);
                    }
        partial[0] = v;
        for (int k = 1; k < nb; k++) partial[k] = want_max ? -DBL_MAX : DBL_MAX;
    });
}

void launch_tree_maxabs(int dim, hipStream_t s, int stride, int nc, const int32_t *leaves,
                        const int32_t *chunk, int n_chunks, const int32_t *tree_chunk, int n_trees,
                        const double *var, double *partial, double *out)
{
    if (n_trees <= 0) return;
    emul_submit(s, [=] {
        const int sp = nc + 2;
        for (int g = 0; g < n_chunks; g++) {
            double v = 0;
            for (int e = chunk[g]; e < chunk[g + 1]; e++)
                for (int k = dim == 3 ? 1 : 0; k <= (dim == 3 ? nc : 0); k++)
                    for (int j = 1; j <= nc; j++)
                        for (int i = 1; i <= nc; i++) {
                            const double x = std::fabs(var[(size_t)leaves[e] * stride + i + sp * (j + sp * k)]);
                            v = x > v ? x : v;
                        }
            partial[g] = v;
        }
        for (int t = 0; t < n_trees; t++) {
            double v = 0;
            for (int q = tree_chunk[t]; q < tree_chunk[t + 1]; q++) v = partial[q] > v ? partial[q] : v;
            out[t] = v;
        }
    });
}

void launch_reduce_partials(hipStream_t s, const double *partial, int n, int want_max, double *out)
{
    emul_submit(s, [=] {
        double v = want_max ? -DBL_MAX : DBL_MAX;
        for (int k = 0; k < n; k++) v = want_max ? (partial[k] > v ? partial[k] : v)
                                                 : (partial[k] < v ? partial[k] : v);
        out[0] = v;
    });
}

int bottom_prepare(int, size_t lds_bytes) { return lds_bytes <= 163840 ? 0 : -1; }

// host stand-in of the device scratch array of the allocation-free entry point
hipError_t scratch_address(int, double **ptr)
{
    static double *p = (double *)emul_alloc(AFG_NOALLOC_MAX_PARTIALS * sizeof(double));
    *ptr = p;
    return p ? hipSuccess : hipErrorMemoryAllocation;
}

template <typename R>
void launch_bottom(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                   const BottomLevels *B, const BottomArgs &A, size_t lds_bytes, int)
{
    emul_submit(s, [=] {
        double *lds = (double *)emul_alloc(lds_bytes ? lds_bytes : 8);
        auto run = [&](auto tau) {
            constexpr int TAU = decltype(tau)::value;
            if (dim == 2) {
                if (V.op_auto)
                    bottom_run<2, 1, R, TAU>(T, V, *B, A, lds, 0, 1, 1, [] {}, [] {}, [](auto v, int) { return v; });
                else
                    bottom_run<2, 0, R, TAU>(T, V, *B, A, lds, 0, 1, 1, [] {}, [] {}, [](auto v, int) { return v; });
            } else {
                if (V.op_auto)
                    bottom_run<3, 1, R, TAU>(T, V, *B, A, lds, 0, 1, 1, [] {}, [] {}, [](auto v, int) { return v; });
                else
                    bottom_run<3, 0, R, TAU>(T, V, *B, A, lds, 0, 1, 1, [] {}, [] {}, [](auto v, int) { return v; });
            }
        };
        if (sizeof(R) == 4 && A.tau64)
            run(std::integral_constant<int, 1>());
        else
            run(std::integral_constant<int, 0>());
        emul_free(lds);
    });
}

template <typename R>
void launch_bottom_residual(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                            const BottomLevels *B, const BottomArgs &A, int)
{
    emul_submit(s, [=] {
        if (dim == 2)
            bottom_residual_run<2, R>(T, V, *B, A, 0, 1);
        else
            bottom_residual_run<3, R>(T, V, *B, A, 0, 1);
    });
}

// The persistent kernel as one workgroup of one thread: the same stage
// sequence (persist_run) with every barrier a no-op, stages in order.
template <typename R>
void launch_persist(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V, const BottomLevels *B,
                    const PersistArgs &P)
{
    emul_submit(s, [=] {
        auto none = [] {};
        auto same = [](auto v, int) { return v; };
        if (dim == 2) {
            if (V.op_auto)
                persist_run<2, 1, R>(T, V, *B, P.A, P.set_residual, 0, 1, 0, 1, 1, none, none, same);
            else
                persist_run<2, 0, R>(T, V, *B, P.A, P.set_residual, 0, 1, 0, 1, 1, none, none, same);
        } else {
            if (V.op_auto)
                persist_run<3, 1, R>(T, V, *B, P.A, P.set_residual, 0, 1, 0, 1, 1, none, none, same);
            else
                persist_run<3, 0, R>(T, V, *B, P.A, P.set_residual, 0, 1, 0, 1, 1, none, none, same);
        }
    });
}

template <typename R>
void launch_residual_levels(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                            const BottomLevels *B, int hl, int64_t)
{
    emul_submit(s, [=] {
        if (dim == 2)
            residual_levels_run<2, R>(T, V, *B, hl, 0, 1);
        else
            residual_levels_run<3, R>(T, V, *B, hl, 0, 1);
    });
}

#define AFG_INSTANTIATE(R)                                                                               \
    template int launch_gsrb_split<R>(int, hipStream_t, const DevTree &, const LevelArgs &,            \
                                      const MgVarsT<R> &, int, int, int, const ListArgs *);                              \
    template int launch_gsrb_pair<R>(int, hipStream_t, const DevTree &, const LevelArgs &,             \
                                     const MgVarsT<R> &, const R *, R *, int, int);                    \
    template void launch_fill_gc<R>(int, hipStream_t, const DevTree &, const LevelArgs &, R *);         \
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
                                            const BottomLevels *, int, int64_t);                       \
    template void launch_persist<R>(int, hipStream_t, const DevTree &, const MgVarsT<R> &,             \
                                    const BottomLevels *, const PersistArgs &);
AFG_INSTANTIATE(double)
AFG_INSTANTIATE(float)

}  // namespace afg
