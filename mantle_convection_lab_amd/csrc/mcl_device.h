// Device utilities shared by the gfx950 kernels of libmcl_hip.so:
//  * a grid-wide barrier for persistent (cooperatively launched) kernels,
//  * compensated FP64 reductions (double-double accumulation, exact products),
//  * the bilinear-quad shape tables (Shape_function.c:17-121 values).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef MCL_BLOCK
#define MCL_BLOCK 512   // threads per workgroup of every kernel (8 waves: one CU each)
#endif
#define MCL_WAVE 64
#define MCL_MAX_LEVELS 20
#define MCL_MAX_WG 2048
// barrier words: gen at 0, abort at 32, arrive[MCL_MAX_WG] from 64
#define MCL_SYNC_WORDS (64 + MCL_MAX_WG)
#define MCL_SPIN_LIMIT (1u << 24)

// Shape tables, uploaded once from the host (computed there with the
// reference's expressions so the values are identical bit for bit).
struct ShapeTables {
    double N[4][4];      // N[a][g] at Gauss point g
    double dxi[4][4];    // dN_a/dxi at g
    double deta[4][4];   // dN_a/deta at g
    double dxic[4], detac[4];
};

// Per-level element sizes (G.dx, G.dz) passed by value to kernels.
struct LevelSizes {
    double dx[MCL_MAX_LEVELS];
    double dz[MCL_MAX_LEVELS];
};

// ---------------------------------------------------------------- grid barrier
// Flag barrier with a master workgroup (no atomic read-modify-write, so the
// cost does not grow with serialised arrivals on one address).  Thread 0 of
// every workgroup publishes the workgroup's writes with an agent-scope release
// fence and then stores the new generation into its own arrive[] slot;
// workgroup 0 polls all arrive[] slots (one slot per thread, coalesced), then
// acquires, releases and stores the generation; the other workgroups poll the
// generation and acquire.  Polls are relaxed agent-scope loads (no cache
// maintenance per poll).  Every thread keeps the generation count itself
// (all threads of the grid start from the same value), so the barrier is
// monotone across launches: a slot or generation counts as reached when it is
// at or past the target.  Spins are bounded: past MCL_SPIN_LIMIT polls the
// barrier sets *abort and returns false, every spinner checks *abort, and
// every caller then leaves its loops, so all waves reach the end of the
// kernel.  After an abort the host zeroes gen, abort and arrive[].
struct GridSync {
    unsigned *gen;
    unsigned *abort;
    unsigned *arrive;   // [MCL_MAX_WG]
    unsigned nblocks;
};

__device__ __forceinline__ void spin_until(const unsigned *flag, unsigned target, unsigned *abort)
{
    unsigned spins = 0;
    while ((int)(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > MCL_SPIN_LIMIT) {
            __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        if ((spins & 1023u) == 0 && __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    }
}

__device__ __forceinline__ bool grid_sync(const GridSync &gs, unsigned &my_gen)
{
    __shared__ unsigned s_abort;
    const unsigned target = my_gen + 1;
    my_gen = target;
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(gs.arrive + blockIdx.x, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blockIdx.x == 0) {
        for (unsigned b = threadIdx.x; b < gs.nblocks; b += blockDim.x) spin_until(gs.arrive + b, target, gs.abort);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(gs.gen, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (threadIdx.x == 0) {
        spin_until(gs.gen, target, gs.abort);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (threadIdx.x == 0) s_abort = __hip_atomic_load(gs.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return s_abort == 0;
}

// Point-to-point wait on the same generation counter: every workgroup
// publishes its arrival as in grid_sync, then waits only for the workgroups of
// its neighbour list (thread t polls the slot of its neighbour nbr, nbr < 0 for
// none).  Used between phases whose data flow is confined to neighbours; the
// counters stay consistent with grid_sync because every workgroup passes the
// same sequence of waits.
__device__ __forceinline__ bool nbr_sync(const GridSync &gs, unsigned &my_gen, int nbr)
{
    __shared__ unsigned s_abort;
    const unsigned target = my_gen + 1;
    my_gen = target;
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(gs.arrive + blockIdx.x, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (nbr >= 0) {
        spin_until(gs.arrive + nbr, target, gs.abort);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (threadIdx.x == 0) s_abort = __hip_atomic_load(gs.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return s_abort == 0;
}

// Neighbour wait for phases whose cross-workgroup vectors are accessed only
// with agent-scope atomic loads and stores (coherent at the device level: a
// store writes through the XCD's L2, a load never returns a stale L2 line).
// Every wave first waits for all its memory operations to complete
// (s_waitcnt 0: its stores are visible device-wide and its loads have
// returned), so the arrival flag is ordered after them without an L2
// write-back; a waiter issues its next loads only after it has seen the flag,
// and they are coherent, so no L2 invalidate is needed either.  The L2 thus
// keeps the read-only matrices across phases.  Plain loads or stores of those
// vectors must stay on the far side of a fenced grid_sync.
__device__ __forceinline__ bool nbr_sync_coherent(const GridSync &gs, unsigned &my_gen, int nbr)
{
    __shared__ unsigned s_abort;
    const unsigned target = my_gen + 1;
    my_gen = target;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(gs.arrive + blockIdx.x, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (nbr >= 0) spin_until(gs.arrive + nbr, target, gs.abort);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x == 0) s_abort = __hip_atomic_load(gs.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return s_abort == 0;
}

// Grid barrier for phases whose cross-workgroup data is accessed only with
// agent-scope atomic loads and stores: the flag protocol of grid_sync with the
// ordering of nbr_sync_coherent (every wave's memory operations complete
// before the arrival; loads after the release are coherent), so no L2
// write-back or invalidate.
__device__ __forceinline__ bool grid_sync_coherent(const GridSync &gs, unsigned &my_gen)
{
    __shared__ unsigned s_abort;
    const unsigned target = my_gen + 1;
    my_gen = target;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(gs.arrive + blockIdx.x, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (blockIdx.x == 0) {
        for (unsigned b = threadIdx.x; b < gs.nblocks; b += blockDim.x) spin_until(gs.arrive + b, target, gs.abort);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(gs.gen, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
        spin_until(gs.gen, target, gs.abort);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x == 0) s_abort = __hip_atomic_load(gs.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return s_abort == 0;
}

// ------------------------------------------------------- compensated reduction
// A double-double value (hi + lo).  Sums use TwoSum, products TwoProduct via
// fma, so a dot product is accumulated as if in ~106-bit precision and rounded
// once at the end (Ogita, Rump, Oishi 2005, "Accurate sum and dot product").
struct dd {
    double hi, lo;
};

__device__ __forceinline__ dd dd_zero() { return dd{0.0, 0.0}; }

__device__ __forceinline__ dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return dd{s, e};
}

__device__ __forceinline__ dd dd_add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    double lo = s.lo + (a.lo + b.lo);
    double hi = s.hi + lo;
    return dd{hi, lo - (hi - s.hi)};
}

__device__ __forceinline__ dd dd_add_prod(dd acc, double x, double y)
{
    double p = x * y;
    double pe = __builtin_fma(x, y, -p);
    dd s = two_sum(acc.hi, p);
    double lo = s.lo + (acc.lo + pe);
    double hi = s.hi + lo;
    return dd{hi, lo - (hi - s.hi)};
}

__device__ __forceinline__ dd dd_shfl_xor(dd v, int mask)
{
    return dd{__shfl_xor(v.hi, mask, MCL_WAVE), __shfl_xor(v.lo, mask, MCL_WAVE)};
}

__device__ __forceinline__ double dmax_shfl(double v)
{
    for (int m = MCL_WAVE / 2; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, MCL_WAVE));
    return v;
}

// Rows of a kernel's reduction scratch (__shared__ dd lds[MCL_RED_ROWS]
// [MCL_BLOCK / MCL_WAVE]): a reduction of NR values uses rows 0 .. NR - 1.
#define MCL_RED_ROWS 8

// Block reduction of NR double-double values (fixed order, deterministic).
template <int NR>
__device__ __forceinline__ void block_reduce_dd(dd (&v)[NR], dd (*lds)[MCL_BLOCK / MCL_WAVE])
{
    static_assert(NR <= MCL_RED_ROWS, "the reduction scratch has MCL_RED_ROWS rows");
    const int lane = threadIdx.x & (MCL_WAVE - 1), wave = threadIdx.x / MCL_WAVE;
#pragma unroll
    for (int r = 0; r < NR; r++)
        for (int m = MCL_WAVE / 2; m > 0; m >>= 1) v[r] = dd_add(v[r], dd_shfl_xor(v[r], m));
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int r = 0; r < NR; r++) lds[r][wave] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NR; r++) {
        dd t = lds[r][0];
        for (int w = 1; w < MCL_BLOCK / MCL_WAVE; w++) t = dd_add(t, lds[r][w]);
        v[r] = t;
    }
}

// Partial-sum exchange through global memory: every workgroup writes its
// partials, and after a grid barrier every workgroup reduces all of them in the
// same order, so every workgroup obtains the identical value.  The partials
// are stored and loaded with agent-scope atomics, so either kind of grid
// barrier (grid_sync or grid_sync_coherent) orders them.
struct RedBuf {
    dd *part;      // [slot][MCL_MAX_WG]
    double *pmax;  // [slot][MCL_MAX_WG]
};

template <int NR>
__device__ __forceinline__ void reduce_publish(const RedBuf &rb, int slot0, dd (&v)[NR],
                                               dd (*lds)[MCL_BLOCK / MCL_WAVE])
{
    block_reduce_dd<NR>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int r = 0; r < NR; r++) {
            dd *p = rb.part + (slot0 + r) * MCL_MAX_WG + blockIdx.x;
            __hip_atomic_store(&p->hi, v[r].hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&p->lo, v[r].lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
}

template <int NR>
__device__ __forceinline__ void reduce_collect(const RedBuf &rb, int slot0, int nblocks, double (&out)[NR],
                                               dd (*lds)[MCL_BLOCK / MCL_WAVE])
{
    dd v[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) v[r] = dd_zero();
    for (int b = threadIdx.x; b < nblocks; b += MCL_BLOCK)
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const dd *p = rb.part + (slot0 + r) * MCL_MAX_WG + b;
            const dd t{__hip_atomic_load(&p->hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __hip_atomic_load(&p->lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
            v[r] = dd_add(v[r], t);
        }
    block_reduce_dd<NR>(v, lds);
#pragma unroll
    for (int r = 0; r < NR; r++) out[r] = v[r].hi + v[r].lo;
}

__device__ __forceinline__ double block_max(double v, double *lds)
{
    const int lane = threadIdx.x & (MCL_WAVE - 1), wave = threadIdx.x / MCL_WAVE;
    v = dmax_shfl(v);
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double t = lds[0];
    for (int w = 1; w < MCL_BLOCK / MCL_WAVE; w++) t = fmax(t, lds[w]);
    return t;
}

__device__ __forceinline__ double block_min(double v, double *lds)
{
    const int lane = threadIdx.x & (MCL_WAVE - 1), wave = threadIdx.x / MCL_WAVE;
    for (int m = MCL_WAVE / 2; m > 0; m >>= 1) v = fmin(v, __shfl_xor(v, m, MCL_WAVE));
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double t = lds[0];
    for (int w = 1; w < MCL_BLOCK / MCL_WAVE; w++) t = fmin(t, lds[w]);
    return t;
}
