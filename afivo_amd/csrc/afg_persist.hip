// afg_persist.hip -- the persistent kernel of the afivo_amd FAS multigrid
// path: the lower part of a V-cycle (the levels with few cells, where a
// launch per stage costs more than the stage's work) in one launch of a few
// workgroups that run on one XCD and step through the stages together.
//
// Workgroups of one XCD share its L2.  Between stages they meet at a barrier
// built on a device-scope counter in that L2; each workgroup first waits for
// its own stores (s_waitcnt), so that what it wrote is in the L2 when it
// arrives.  The stages read the arrays they exchange (phi, rhs, tmp) with
// device-scope loads (AFG_COHERENT_LOADS), which miss in the compute unit's
// vector L1 -- that cache is not kept coherent with other compute units'
// stores -- and read the shared L2.  Every wait is bounded in time: a
// workgroup that waits longer than PERSIST_SPIN_TICKS sets the error flag,
// every wait then ends at once, the kernel runs to its end and the host
// reports the error (the results are then invalid).
#define AFG_COHERENT_LOADS 1

#include "afivo_amd.h"
#include "afg_internal.h"
#include "afg_bodies.h"
#include "afg_lane.h"

namespace afg {

namespace {

// bound of one barrier wait in s_memrealtime ticks (100 MHz): 200 ms
constexpr unsigned long long PERSIST_SPIN_TICKS = 20000000ull;

template <int DIM, int AUTO, typename R>
__global__ __launch_bounds__(PERSIST_BLOCK) void k_persist(DevTree T, MgVarsT<R> V,
                                                           const BottomLevels *__restrict__ B, PersistArgs P)
{
    if ((int)(blockIdx.x & 7) != P.xcd) return;
    const int rank = (int)(blockIdx.x >> 3);
    unsigned *const arrive = P.sync, *const depart = P.sync + 1, *const err = P.sync + 2;
    unsigned target = 0;
    int n_trace = 0;
    auto stamp = [&] {
        if (P.trace && rank == 0 && threadIdx.x == 0 && n_trace < P.trace_cap)
            P.trace[n_trace++] = __builtin_amdgcn_s_memrealtime();
    };
    stamp();
    auto gsync = [&] {
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's stores have reached the L2
        __syncthreads();
        target += (unsigned)P.n_rank;
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
            for (unsigned it = 1; __hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target;
                 it++) {
                if ((it & 63) == 0 && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                                       __builtin_amdgcn_s_memrealtime() - t0 > PERSIST_SPIN_TICKS)) {
                    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        stamp();
    };
    auto wgsync = [] {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };
    // every workgroup must share the L2: each publishes the XCD it runs on
    // and checks the others' (error flag 2 when they differ)
    unsigned xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    xcc &= 0xf;
    if (threadIdx.x == 0) __hip_atomic_store(P.sync + 4 + rank, xcc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    gsync();
    if (threadIdx.x < (unsigned)P.n_rank &&
        __hip_atomic_load(P.sync + 4 + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != xcc + 1)
        __hip_atomic_store(err, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    persist_run<DIM, AUTO, R>(T, V, *B, P.A, P.set_residual, rank, P.n_rank, threadIdx.x, blockDim.x, warpSize,
                              gsync, wgsync, [](auto x, int mask) { return lane_xor(x, mask); });
    // the last workgroup to leave clears the counters and the XCD slots for
    // the next launch: every other one has passed its last wait (and read the
    // slots) when it counts its departure
    if (threadIdx.x == 0 &&
        __hip_atomic_fetch_add(depart, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (unsigned)P.n_rank - 1) {
        __hip_atomic_store(arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(depart, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int r = 0; r < P.n_rank; r++) __hip_atomic_store(P.sync + 4 + r, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

template <typename R>
void launch_persist(int dim, hipStream_t s, const DevTree &T, const MgVarsT<R> &V, const BottomLevels *B,
                    const PersistArgs &P)
{
    const dim3 grid(8 * P.n_rank), block(PERSIST_BLOCK);
    if (dim == 2) {
        if (V.op_auto)
            k_persist<2, 1, R><<<grid, block, 0, s>>>(T, V, B, P);
        else
            k_persist<2, 0, R><<<grid, block, 0, s>>>(T, V, B, P);
    } else {
        if (V.op_auto)
            k_persist<3, 1, R><<<grid, block, 0, s>>>(T, V, B, P);
        else
            k_persist<3, 0, R><<<grid, block, 0, s>>>(T, V, B, P);
    }
}

template void launch_persist<double>(int, hipStream_t, const DevTree &, const MgVarsT<double> &,
                                     const BottomLevels *, const PersistArgs &);
template void launch_persist<float>(int, hipStream_t, const DevTree &, const MgVarsT<float> &,
                                    const BottomLevels *, const PersistArgs &);

}  // namespace afg
