// afg_internal.h -- launchers of the HIP kernels, called by the host
// orchestration of the afivo_amd multigrid library.
#pragma once

#include <hip/hip_runtime.h>

#include "afg_types.h"

namespace afg {

// ---- launchers (afg_kernels.hip) ----
// R, the value type: double, or float for the FP32 stages of the
// mixed-precision solve (both are instantiated).
template <typename R>
void launch_fill_gc(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, R *var);
// ghost fill of L and tmp = phi - tmp on the parents P of the level below
template <typename R>
void launch_fill_diff(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const ListArgs &P, const MgVarsT<R> &V);
template <typename R>
void launch_diff(int dim, hipStream_t s, const DevTree &T, const ListArgs &P, const MgVarsT<R> &V);
template <typename R>
void launch_gsrb_pass(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const MgVarsT<R> &V, int rb, int fly);
// passes rbA and rbA + 1 in one launch (nc 8 and 16): the level's phi read
// from `in` and written to `out` (V.phi holds the other levels); returns 0
// (nothing launched) for other cell counts
template <typename R>
int launch_gsrb_pair(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V,
                     const R *in, R *out, int rbA, int fly);
// one colour pass of a relaxation block on the colour-split copy V.phi_rb
// (3D, nc 8 and 16, tiles): mode 0 the first pass of the block, 1 a middle
// one, 2 the last, 3 the last followed by the level's ghost fill and, with P,
// tmp = phi - tmp over the parents P; returns 0 (nothing launched) where it
// does not apply
template <typename R>
int launch_gsrb_split(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<R> &V, int rb,
                      int mode, int fly, const ListArgs *P = nullptr);
template <typename R>
void launch_residual(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const MgVarsT<R> &V);
// k_residual over a level (FP64) that also writes (float) of the residual
// into r32 (the FP32 rhs pool) and r32_rb (its colour-split copy, or null)
void launch_residual32(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const MgVarsT<double> &V,
                       float *r32, float *r32_rb, int rb_stride);
// from_rb: the level's phi is in the colour-split copy V.phi_rb (3D tiles)
template <typename R>
void launch_uc_child(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const MgVarsT<R> &V, int keep_child_tmp, int fly, int from_rb = 0);
// L: the parent level (all its boxes get their ghosts filled)
template <typename R>
void launch_uc_parent(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                      const MgVarsT<R> &V, int copy_tmp);
// to_rb: the level's phi is in the colour-split copy V.phi_rb
template <typename R>
void launch_correct(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                    const MgVarsT<R> &V, int fly_diff, int to_rb = 0);
// rhs by colour (MgVarsT::rhs_rb) of every box of a level, from rhs
template <typename R>
void launch_split_rhs(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const R *rhs, R *rhs_rb,
                      int rb_stride);
template <typename R>
void launch_restrict(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const R *from, R *to, const R *lsf = nullptr, int rstr_mode = 0);
template <typename R>
void launch_prolong2(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const R *from, R *to, int add);
template <typename R>
void launch_fill_value(int dim, hipStream_t s, const DevTree &T, int slot0, int n_slots,
                       int nc, R *var, double value);
template <typename R>
void launch_copy_boxes(int dim, hipStream_t s, const DevTree &T, int slot0, int n_slots, int nc,
                       const R *from, R *to);
// tree update: to[slot] = from[src[slot]] (whole box arrays of `stride`
// values), zeros where src[slot] < 0
void launch_gather(hipStream_t s, int stride, int n_slots, const int32_t *src, const double *from,
                   double *to);
// a$D_prolong0/1/2 (order 0/1/2) of var from the parents of the n listed
// slots (all with nc cells, their parents with pnc) into their interiors
void launch_prolong_list(int dim, hipStream_t s, const DevTree &T, int nc, int pnc, const int32_t *slots,
                         int n, int order, double *var);
// a$D_prolong0/1/2 (order 0/1/2) of `from` into every box of level L
// (`to`), added to the old values when add
void launch_prolong_level(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const double *from,
                          double *to, int order, int add);
// face ghosts of every box of level L with refinement-boundary method rb
// (RB_*) and boundary condition bc (0: the table; -1 / -2: Dirichlet /
// Neumann zero)
void launch_gc(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, double *var, int rb, int bc);
// the second ghost layer of every box of L (gc2_body) into out, gstride values per slot
void launch_gc2(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L, const double *var, double *out,
                int gstride);
// a$D_consistent_fluxes for the n records of one level (face pool fc); on
// a cylindrical tree cyl_rmin holds r_min(1) per slot (else nullptr) and dr
// is the level's cell size
void launch_flux(int dim, hipStream_t s, const FluxRec *recs, int n, int nc, int fstride, double *fc,
                 const double *cyl_rmin, double dr);
// y = y + x, y - x, a y or a x + b y (ARITH_*) over n values
void launch_arith(hipStream_t s, long long n, int op, double a, double b, const double *x, double *y);
// n values of a pool to FP32, and back (optionally added to the FP64 values)
void launch_to_fp32(hipStream_t s, long long n, const double *from, float *to);
void launch_from_fp32(hipStream_t s, long long n, const float *from, double *to, int add);
void launch_set_tags(int dim, hipStream_t s, const DevTree &T, const LevelArgs &L,
                     const double *eps, const double *lsf);
// Extreme (max or min) of the interiors of the boxes of a slot list: one
// partial per block into `partial`; returns the number of partials written.
int extreme_blocks(int dim, const ListArgs &P);
void launch_extreme_list(int dim, hipStream_t s, const DevTree &T, const ListArgs &P,
                         const double *var, int want_max, double *partial);
// out[0] = extreme of partial[0..n).
void launch_tree_maxabs(int dim, hipStream_t s, int stride, int nc, const int32_t *leaves,
                        const int32_t *chunk, int n_chunks, const int32_t *tree_chunk, int n_trees,
                        const double *var, double *partial, double *out);
void launch_reduce_partials(hipStream_t s, const double *partial, int n, int want_max,
                            double *out);
// The device address of the allocation-free entry points' scratch array
// (AFG_NOALLOC_MAX_PARTIALS doubles) on `device`, which must be the current
// device: looked up once per device (hipGetSymbolAddress) and kept.
hipError_t scratch_address(int device, double **ptr);
// bottom kernel: allow lds_bytes of dynamic LDS (0 on success), then launch
int bottom_prepare(int dim, size_t lds_bytes);
template <typename R>
void launch_bottom(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                   const BottomLevels *B, const BottomArgs &A, size_t lds_bytes, int threads);
template <typename R>
void launch_bottom_residual(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                            const BottomLevels *B, const BottomArgs &A, int threads);
// the residual of levels 0..hl of table B over the HBM pools in one launch
// (max_cells: the most cells any of those levels holds)
template <typename R>
void launch_residual_levels(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V,
                            const BottomLevels *B, int hl, int64_t max_cells);
// ---- the persistent kernel (afg_persist.hip) ----
template <typename R>
void launch_persist(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V, const BottomLevels *B,
                    const PersistArgs &P);

}  // namespace afg
