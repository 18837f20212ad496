// mdrl_device.h -- device helpers shared by the kernels of libmdrl (pair geometry, table
// lookup, fixed-point conversion, wave-level compaction, counting-sort increments).
//
// Built with -ffp-contract=off: every rounding happens where it is written, fused
// multiply-adds are explicit fmaf()/fma() calls, so the oracle can restate the arithmetic
// bit for bit.
//
// Reference paths: D/ = Range_Limited_demo/SourceCode/, R/ = Range_Limited/,
// M/ = D/MatlabScripts/, O/ = D/OpenCL_HDL_Lib/.
#pragma once
#include <hip/hip_runtime.h>

#include "mdrl_internal.h"

namespace mdrl {

// ---------------------------------------------------------------------------------
// shared device helpers
// ---------------------------------------------------------------------------------

// Table entry of r2: {exp8(r2) - offset, mantissa[22:15]}, segment 0 when out of range.
// D/RL_LJ_Evaluate_Pairs_1st_Order.v:176-199.  Bits 30..15 of r2 are exp8 << 8 | bin, so
// the entry is those bits less offset << 8 when that lies below nseg << 8 (one unsigned
// compare covers both ends), else the bin alone.
__device__ __forceinline__ unsigned lut_index(float r2, int off, int nseg)
{
    const unsigned raw = (__float_as_uint(r2) >> 15) & 0xffffu;
    const unsigned rel = raw - ((unsigned)off << 8);
    return rel < ((unsigned)nseg << 8) ? rel : (raw & 0xffu);
}

// rint(f * 2^32) as int64, saturated at +-2^62 (flag bit 0 set on saturation or NaN).
// Fast path: for |f*2^32| < 2^51 the sum f*2^32 + 1.5*2^52 (one rounding, the fma) is an
// integer whose low mantissa bits are the two's-complement result.
__device__ __forceinline__ long long fix_of(float f, unsigned &flag)
{
    double d = (double)f;
    double t = fma(d, 4294967296.0, 6755399441055744.0);
    if (fabs(d) < 524288.0) {
        return __double_as_longlong(t) - 0x4338000000000000LL;
    }
    double x = d * 4294967296.0;
    if (x != x) { flag |= 2u; return 0; }
    if (x > 0x1p62) { x = 0x1p62; flag |= 1u; }
    if (x < -0x1p62) { x = -0x1p62; flag |= 1u; }
    return (long long)rint(x);
}

// fix_of of the three components of a pair force, with one range test for the fast path
// (|f| summed, so a NaN or an infinity takes the general path).
__device__ __forceinline__ void fix3(float fx, float fy, float fz, long long &kx, long long &ky,
                                     long long &kz, unsigned &flag)
{
    const float m = (fabsf(fx) + fabsf(fy)) + fabsf(fz);
    if (m < 524288.0f) {
        kx = __double_as_longlong(fma((double)fx, 4294967296.0, 6755399441055744.0)) - 0x4338000000000000LL;
        ky = __double_as_longlong(fma((double)fy, 4294967296.0, 6755399441055744.0)) - 0x4338000000000000LL;
        kz = __double_as_longlong(fma((double)fz, 4294967296.0, 6755399441055744.0)) - 0x4338000000000000LL;
    } else {
        kx = fix_of(fx, flag);
        ky = fix_of(fy, flag);
        kz = fix_of(fz, flag);
    }
}

// fix3 of (fx, fy, fz) given as fx * 2^32 etc. (the products of a force scaled by 2^32,
// which carry the same bits times 2^32): the fast path is one FP64 add of 1.5 * 2^52
// (rounding to the integer, ties to even) and the integer's bits; saturation at +-2^62
// and NaN as fix_of.
__device__ __forceinline__ long long fix_scaled_slow(float x, unsigned &flag)
{
    double d = (double)x;
    if (d != d) { flag |= 2u; return 0; }
    if (d > 0x1p62) { d = 0x1p62; flag |= 1u; }
    if (d < -0x1p62) { d = -0x1p62; flag |= 1u; }
    return (long long)rint(d);
}
__device__ __forceinline__ void fix3_scaled(float x, float y, float z, long long &kx, long long &ky,
                                            long long &kz, unsigned &flag)
{
    const float m = (fabsf(x) + fabsf(y)) + fabsf(z);
    if (m < 0x1p51f) {
        kx = __double_as_longlong((double)x + 6755399441055744.0) - 0x4338000000000000LL;
        ky = __double_as_longlong((double)y + 6755399441055744.0) - 0x4338000000000000LL;
        kz = __double_as_longlong((double)z + 6755399441055744.0) - 0x4338000000000000LL;
    } else {
        kx = fix_scaled_slow(x, flag);
        ky = fix_scaled_slow(y, flag);
        kz = fix_scaled_slow(z, flag);
    }
}

__device__ __forceinline__ unsigned lane_rank(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Counting-sort increments with one atomic per run of equal keys in consecutive lanes:
// slots stay cell-sorted between steps, so a wave's atoms mostly share one or two cells.
// Active lanes must form a prefix of the wave.  Returns, for each active lane, the old
// value of count[key] plus the lane's rank within its run (its slot for a scatter).
__device__ __forceinline__ int run_add(int *count, int key, bool active)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1);
    const unsigned long long act = __ballot(active);
    const int nact = __popcll(act);
    const bool head = active && (lane == 0 || prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    int base = 0;
    if (head) {
        const unsigned long long after = heads & ~upto;
        const int end = after ? __ffsll((long long)after) - 1 : nact;
        base = atomicAdd(&count[key], end - lane);
    }
    const int hl = 63 - __clzll(heads & upto);
    const int hb = __shfl(base, active ? hl : lane);
    return hb + (lane - hl);
}

// Which of the kCountCopies copies of the per-cell counters this block adds into: blocks
// are dealt round-robin to the XCDs, so each copy is touched from one XCD, and a cell's
// increments spread over kCountCopies addresses instead of one.
__device__ __forceinline__ int count_copy() { return (int)(blockIdx.x % kCountCopies); }

template <typename T>
__device__ __forceinline__ void atomic_add_agent(T *p, T v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void atomic_add_wg(T *p, T v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void atomic_sub_wg(T *p, T v)
{
    __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Workgroup barrier that orders LDS only: global loads, stores and atomics issued before
// it stay in flight (a full __syncthreads() would wait for every one of them).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Cell of one coordinate: (int)(x * inv_edge) clamped to [0, nc-1]; the clamp comes
// first so a non-finite coordinate maps to a valid cell (NaN -> 0).
__device__ __forceinline__ int cell_of(float x, float inv_edge, int nc)
{
    float t = x * inv_edge;
    t = fminf(fmaxf(t, 0.0f), (float)(nc - 1));
    return (int)t;
}

// Half-shell neighbour offsets: home 222 then {223,231,232,233,311,312,313,321,322,323,
// 331,332,333} (D/Particle_Pair_Gen_HalfShell_64_Cells.v:213-226).
static __constant__ int c_hs[14][3] = {
    {0, 0, 0},   {0, 0, 1},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1},
    {1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0},
    {1, 0, 1},   {1, 1, -1}, {1, 1, 0},  {1, 1, 1},
};

// ---------------------------------------------------------------------------------
// pair geometry and pair force (shared by the force and energy kernels)
// ---------------------------------------------------------------------------------

// PAR: dx = (xi - xj) - S (the reference's subtract-then-image order, r2 unfused);
// otherwise xj arrives pre-shifted (xj + S rounded once at staging) and r2 uses fma.
template <bool PAR>
__device__ __forceinline__ float pair_geom(float4 pi, float4 pj, float sx, float sy, float sz,
                                           float &dx, float &dy, float &dz)
{
    if (PAR) {
        dx = pi.x - pj.x; dx = dx - sx;
        dy = pi.y - pj.y; dy = dy - sy;
        dz = pi.z - pj.z; dz = dz - sz;
        float r2 = dx * dx + dy * dy;
        return r2 + dz * dz;
    } else {
        dx = pi.x - pj.x;
        dy = pi.y - pj.y;
        dz = pi.z - pj.z;
        return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    }
}

// Table coefficients of r2 (c0, c1 of r^-14, r^-8, r^-3) from the global table (6 floats
// per entry), entry {exp8(r2) - offset, mantissa[22:15]} as the reference addresses it.
// The reads go through the vector L1/L2 rather than an LDS copy: the force kernel's LDS is
// busy with its tiles, queues and accumulator atomics, and random table gathers there cost
// bank conflicts (92k system: 0.181 ms per force launch with a window of the table in LDS,
// 0.160 ms with these reads).
template <bool COUL>
__device__ __forceinline__ void table_coeffs(float r2, const float *__restrict__ g_tab, int off, int nseg,
                                             float c[6])
{
    // (a 32-bit byte offset from the table's base, so the loads take the scalar-base form)
    const float *q = reinterpret_cast<const float *>(reinterpret_cast<const char *>(g_tab) +
                                                     __umul24(lut_index(r2, off, nseg), 24u));
    const float2 u = *reinterpret_cast<const float2 *>(q);
    const float2 v = *reinterpret_cast<const float2 *>(q + 2);
    c[0] = u.x; c[1] = u.y; c[2] = v.x; c[3] = v.y;
    if (COUL) {
        const float2 w = *reinterpret_cast<const float2 *>(q + 4);
        c[4] = w.x; c[5] = w.y;
    } else {
        c[4] = 0.f; c[5] = 0.f;
    }
}

// Scalar pair force F (kJ/mol/nm per A of separation): table path
// D/RL_LJ_Evaluate_Pairs_1st_Order.v:409-475 with the per-pair multipliers of
// D/RL_Evaluate_Pairs_1st_Order.v:53-55 (A = 10 eps sig^12, B = 10 eps sig^6, qq = qs_i qs_j).
template <bool COUL, bool PAR, bool ANA>
__device__ __forceinline__ float pair_scalar(float r2, float A, float B, float qq,
                                             const float *__restrict__ g_tab, int off, int nseg)
{
    if (ANA) {
        float ir2 = 1.0f / r2;
        float ir6 = ir2 * ir2;
        ir6 = ir6 * ir2;
        float ir8 = ir6 * ir2;
        float ir14 = ir8 * ir6;
        float F = fmaf(A, ir14, -(B * ir8));
        if (COUL) {
            float ir3 = ir2 * sqrtf(ir2);
            ir3 = ir3 * 0.0796178343949044585987261f;
            F = fmaf(qq, ir3, F);
        }
        return F;
    }
    float c[6];
    table_coeffs<COUL>(r2, g_tab, off, nseg, c);
    float F;
    if (PAR) {
        float t14 = c[1] * r2; t14 = t14 + c[0];
        float t8 = c[3] * r2;  t8 = t8 + c[2];
        float x = A * t14, y = B * t8;
        F = x - y;
        if (COUL) {
            float t3 = c[5] * r2; t3 = t3 + c[4];
            float z = qq * t3;
            F = F + z;
        }
    } else {
        float t14 = fmaf(c[1], r2, c[0]);
        float t8 = fmaf(c[3], r2, c[2]);
        F = fmaf(A, t14, -(B * t8));
        if (COUL) {
            float t3 = fmaf(c[5], r2, c[4]);
            F = fmaf(qq, t3, F);
        }
    }
    return F;
}


}  // namespace mdrl
