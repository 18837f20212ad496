// mdrl_kernels.hip -- gfx950 kernels of the range-limited force path and their launchers.
//
// Built with -ffp-contract=off: every rounding happens where it is written, fused
// multiply-adds are explicit fmaf()/fma() calls, so the oracle can restate the
// arithmetic bit for bit.
//
// Reference paths: D/ = Range_Limited_demo/SourceCode/, R/ = Range_Limited/,
// M/ = D/MatlabScripts/, O/ = D/OpenCL_HDL_Lib/.
#include <hip/hip_runtime.h>

#include "mdrl_internal.h"

namespace mdrl {

// ---------------------------------------------------------------------------------
// shared device helpers
// ---------------------------------------------------------------------------------

// Table entry of r2: {exp8(r2) - offset, mantissa[22:15]}, segment 0 when out of range.
// D/RL_LJ_Evaluate_Pairs_1st_Order.v:176-199.
__device__ __forceinline__ int lut_seg(float r2, int off, int nseg)
{
    int e = (int)((__float_as_uint(r2) >> 23) & 0xffu);
    int seg = e - off;
    return (seg < 0 || seg >= nseg) ? 0 : seg;
}
__device__ __forceinline__ int lut_bin(float r2) { return (int)((__float_as_uint(r2) >> 15) & 0xffu); }

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
__constant__ int c_hs[14][3] = {
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

// Table coefficients of r2 (c0, c1 of r^-14, r^-8, r^-3).  Entries [win_lo, win_lo + nwin)
// -- the bins reachable below rc2 down to a floor seven octaves lower -- are read from
// LDS; the rare r2 outside (overlapping atoms, or the reference's segment-0 fallback)
// re-reads its entry from the global copy.
template <bool COUL, int STR>
__device__ __forceinline__ void table_coeffs(float r2, const float *s_tab, const float *__restrict__ g_tab,
                                             int off, int nseg, int win_lo, int nwin, float c[6])
{
    const int seg = lut_seg(r2, off, nseg);
    const int idx = seg * 256 + lut_bin(r2);
    const bool inw = (unsigned)(idx - win_lo) < (unsigned)nwin;
    const int li = inw ? idx - win_lo : 0;
    if (STR == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(s_tab + li * 4);
        c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w; c[4] = 0.f; c[5] = 0.f;
    } else {
        const float2 u = *reinterpret_cast<const float2 *>(s_tab + li * 6);
        const float2 v = *reinterpret_cast<const float2 *>(s_tab + li * 6 + 2);
        const float2 w = *reinterpret_cast<const float2 *>(s_tab + li * 6 + 4);
        c[0] = u.x; c[1] = u.y; c[2] = v.x; c[3] = v.y; c[4] = w.x; c[5] = w.y;
    }
    if (!inw) {
        const float *q = g_tab + (size_t)idx * 6;
        c[0] = q[0]; c[1] = q[1]; c[2] = q[2]; c[3] = q[3];
        if (COUL) { c[4] = q[4]; c[5] = q[5]; }
    }
}

// Scalar pair force F (kJ/mol/nm per A of separation): table path
// D/RL_LJ_Evaluate_Pairs_1st_Order.v:409-475 with the per-pair multipliers of
// D/RL_Evaluate_Pairs_1st_Order.v:53-55 (A = 10 eps sig^12, B = 10 eps sig^6, qq = qs_i qs_j).
template <bool COUL, bool PAR, bool ANA, int STR>
__device__ __forceinline__ float pair_scalar(float r2, float A, float B, float qq,
                                             const float *s_tab, const float *__restrict__ g_tab,
                                             int off, int nseg, int win_lo, int nwin)
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
    table_coeffs<COUL, STR>(r2, s_tab, g_tab, off, nseg, win_lo, nwin, c);
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

// ---------------------------------------------------------------------------------
// k_rl_halfshell: one workgroup of 8 waves per (home cell, group of half-shell offsets).
//
// The home cell's atoms are staged in LDS (tile of TI = 256), each neighbour cell of the
// group is staged (pre-shifted by its periodic image) as a tile of TJ.  Wave w streams
// neighbours w, w+8, w+16, ... of the tile, one per step as an LDS broadcast, against
// every 64-atom chunk of the home tile (one home atom per lane, up to four chunks held in
// registers, so one neighbour read feeds four independent tests).  The r2 <= rc2 test of
// D/Filter_Logic.v:234 is a ballot; passing (i, j) pairs are compacted into a per-wave LDS
// ring (the filter buffers + arbiter of D/Filter_Bank.v:548-606) and, whenever 64 are
// queued, all 64 lanes evaluate them densely: table force from an LDS window, then the
// reference particle gets +F and the neighbour -F (D/RL_LJ_Evaluation_Unit.v:137) in LDS
// accumulators, int64 fixed point (exact, order-free) or FP32.  Neighbour forces are
// flushed to HBM with one atomic per touched particle and component after each
// neighbour tile (D/Force_Write_Back_Controller.v:370-429), home forces after the group's
// last offset (D/Partial_Force_Acc.v:65-129).
// ---------------------------------------------------------------------------------

constexpr int TI = 256, TJ = 256, QCAP = 128, NWAVE = 8, NTHR = 512, NCHUNK = TI / 64;

struct ForceArgs {
    const float4 *pos;
    const int *type;
    const int *cell_start;
    int nc0, nc1, nc2;
    float box0, box1, box2;
    float rc2;
    const float *tab;
    int off, nseg, win_lo, nwin;
    const float2 *pair_ab;
    int ntypes;
    long long *fix;
    float *f32;
    unsigned long long *counters;
    int ngroups;
    int gstart[8];
};

template <bool COUL, bool F32, bool PAR, bool ANA>
__global__ __launch_bounds__(NTHR) void k_rl_halfshell(ForceArgs a)
{
    constexpr int STR = COUL ? 6 : 4;
    extern __shared__ float s_tab[];
    __shared__ float4 s_i[TI];
    __shared__ float4 s_j[TJ];
    __shared__ int s_it[TI];
    __shared__ int s_jt[TJ];
    __shared__ unsigned long long s_iacc[3 * TI];
    __shared__ unsigned long long s_jacc[3 * TJ];
    __shared__ unsigned s_q[NWAVE][QCAP];
    __shared__ float2 s_ab[64];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
    const int task = blockIdx.x;
    const int h = task / a.ngroups, grp = task - h * a.ngroups;
    const int hz = h % a.nc2, hy = (h / a.nc2) % a.nc1, hx = h / (a.nc1 * a.nc2);
    const int hstart = a.cell_start[h], hcount = a.cell_start[h + 1] - hstart;
    if (hcount == 0) return;

    // table window and type-pair multipliers
    if (!ANA) {
        const float *g = a.tab + (size_t)a.win_lo * 6;
        for (int e = tid; e < a.nwin; e += NTHR) {
            const float *p = g + (size_t)e * 6;
            float *q = s_tab + e * STR;
            q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; q[3] = p[3];
            if (STR == 6) { q[4] = p[4]; q[5] = p[5]; }
        }
    }
    for (int e = tid; e < a.ntypes * a.ntypes; e += NTHR) s_ab[e] = a.pair_ab[e];

    float *s_iacc_f = reinterpret_cast<float *>(s_iacc);
    float *s_jacc_f = reinterpret_cast<float *>(s_jacc);
    unsigned *q = s_q[wave];
    unsigned flag = 0;
    unsigned long long npairs = 0, ncand = 0;
    const float rc2 = a.rc2;

    for (int itile = 0; itile < hcount; itile += TI) {
        const int mi = min(TI, hcount - itile);
        const int nch = (mi + 63) >> 6;
        __syncthreads();
        for (int k = tid; k < mi; k += NTHR) {
            s_i[k] = a.pos[hstart + itile + k];
            s_it[k] = a.type[hstart + itile + k];
        }
        for (int k = tid; k < 3 * TI; k += NTHR) s_iacc[k] = 0ull;
        __syncthreads();
        // this lane's home atoms, one per chunk; lanes past the tile hold a far-away
        // point (r2 = inf never passes)
        float4 pi[NCHUNK];
#pragma unroll
        for (int ch = 0; ch < NCHUNK; ++ch) {
            const int ia = ch * 64 + lane;
            pi[ch] = (ia < mi) ? s_i[ia] : make_float4(1e30f, 1e30f, 1e30f, 0.0f);
        }

        for (int o = a.gstart[grp]; o < a.gstart[grp + 1]; ++o) {
            const bool self = (o == 0);
            int rx = hx + c_hs[o][0], ry = hy + c_hs[o][1], rz = hz + c_hs[o][2];
            float sx = 0.f, sy = 0.f, sz = 0.f;
            if (rx < 0) { rx += a.nc0; sx = -a.box0; } else if (rx >= a.nc0) { rx -= a.nc0; sx = a.box0; }
            if (ry < 0) { ry += a.nc1; sy = -a.box1; } else if (ry >= a.nc1) { ry -= a.nc1; sy = a.box1; }
            if (rz < 0) { rz += a.nc2; sz = -a.box2; } else if (rz >= a.nc2) { rz -= a.nc2; sz = a.box2; }
            const int c = (rx * a.nc1 + ry) * a.nc2 + rz;
            const int cstart = a.cell_start[c], ccount = a.cell_start[c + 1] - cstart;

            for (int jtile = 0; jtile < ccount; jtile += TJ) {
                if (self && jtile + TJ <= itile) continue;   // all pairs below the diagonal
                const int mj = min(TJ, ccount - jtile);
                __syncthreads();
                for (int k = tid; k < mj; k += NTHR) {
                    float4 p = a.pos[cstart + jtile + k];
                    if (!PAR) { p.x = p.x + sx; p.y = p.y + sy; p.z = p.z + sz; }
                    s_j[k] = p;
                    s_jt[k] = a.type[cstart + jtile + k];
                }
                for (int k = tid; k < 3 * mj; k += NTHR) s_jacc[k] = 0ull;
                __syncthreads();

                auto eval = [&](unsigned e) {
                    const int ia = (int)(e & 0xffffu), jb = (int)(e >> 16);
                    const float4 pa = s_i[ia], pb = s_j[jb];
                    float dx, dy, dz;
                    const float r2 = pair_geom<PAR>(pa, pb, sx, sy, sz, dx, dy, dz);
                    const float2 ab = s_ab[s_it[ia] * a.ntypes + s_jt[jb]];
                    const float qq = pa.w * pb.w;
                    const float F = pair_scalar<COUL, PAR, ANA, STR>(r2, ab.x, ab.y, qq, s_tab, a.tab,
                                                                     a.off, a.nseg, a.win_lo, a.nwin);
                    const float fx = F * dx, fy = F * dy, fz = F * dz;
                    if (F32) {
                        atomic_add_wg(&s_iacc_f[3 * ia + 0], fx);
                        atomic_add_wg(&s_iacc_f[3 * ia + 1], fy);
                        atomic_add_wg(&s_iacc_f[3 * ia + 2], fz);
                        atomic_add_wg(&s_jacc_f[3 * jb + 0], -fx);
                        atomic_add_wg(&s_jacc_f[3 * jb + 1], -fy);
                        atomic_add_wg(&s_jacc_f[3 * jb + 2], -fz);
                    } else {
                        const long long kx = fix_of(fx, flag), ky = fix_of(fy, flag), kz = fix_of(fz, flag);
                        atomic_add_wg(&s_iacc[3 * ia + 0], (unsigned long long)kx);
                        atomic_add_wg(&s_iacc[3 * ia + 1], (unsigned long long)ky);
                        atomic_add_wg(&s_iacc[3 * ia + 2], (unsigned long long)kz);
                        atomic_sub_wg(&s_jacc[3 * jb + 0], (unsigned long long)kx);
                        atomic_sub_wg(&s_jacc[3 * jb + 1], (unsigned long long)ky);
                        atomic_sub_wg(&s_jacc[3 * jb + 2], (unsigned long long)kz);
                    }
                };

                unsigned head = 0, tail = 0;
                // Rotation: at step sft, lane l tests neighbour (l + sft) mod mj against each
                // of its home atoms, so within a step both the home and the neighbour indices
                // of the queued pairs are distinct (conflict-free accumulator updates) and one
                // LDS read of the neighbour feeds every chunk.  Steps are strided over waves.
                const int base = lane % mj;
                // home-cell pairs are taken once, as (lower, higher) in-cell index: a pair
                // passes only if its neighbour index exceeds lim (lim = -1 off the diagonal)
                int lim[NCHUNK];
#pragma unroll
                for (int ch = 0; ch < NCHUNK; ++ch) lim[ch] = self ? itile + ch * 64 + lane - jtile : -1;
                for (int sft = wave; sft < mj; sft += NWAVE) {
                    int jb = base + sft;
                    jb = (jb >= mj) ? jb - mj : jb;
                    const float4 pj = s_j[jb];
#pragma unroll
                    for (int ch = 0; ch < NCHUNK; ++ch) {
                        if (ch < nch) {
                            float dx, dy, dz;
                            const float r2 = pair_geom<PAR>(pi[ch], pj, sx, sy, sz, dx, dy, dz);
                            const int ia = ch * 64 + lane;
                            const bool pass = (r2 > 0.0f) && (r2 <= rc2) && (jb > lim[ch]);
                            const unsigned long long m = __builtin_amdgcn_ballot_w64(pass);
                            if (pass) q[(tail + lane_rank(m)) & (QCAP - 1)] = (unsigned)ia | ((unsigned)jb << 16);
                            tail += (unsigned)__popcll(m);
                            if (tail - head >= 64u) {
                                __builtin_amdgcn_wave_barrier();
                                eval(q[(head + lane) & (QCAP - 1)]);
                                __builtin_amdgcn_wave_barrier();
                                head += 64u;
                            }
                        }
                    }
                }
                ncand += (unsigned long long)((mj - wave + NWAVE - 1) / NWAVE) * (unsigned long long)nch;
                npairs += tail;
                if (tail != head) {
                    __builtin_amdgcn_wave_barrier();
                    if ((unsigned)lane < tail - head) eval(q[(head + lane) & (QCAP - 1)]);
                    __builtin_amdgcn_wave_barrier();
                }
                __syncthreads();
                // flush neighbour partial forces (D/Force_Write_Back_Controller.v:370-429)
                if (F32) {
                    for (int k = tid; k < 3 * mj; k += NTHR) {
                        float v = s_jacc_f[k];
                        if (v != 0.0f) atomic_add_agent(&a.f32[4 * (size_t)(cstart + jtile + k / 3) + k % 3], v);
                    }
                } else {
                    long long *dst = a.fix + 3 * (size_t)(cstart + jtile);
                    for (int k = tid; k < 3 * mj; k += NTHR) {
                        unsigned long long v = s_jacc[k];
                        if (v) atomic_add_agent(reinterpret_cast<unsigned long long *>(dst + k), v);
                    }
                }
            }
        }
        __syncthreads();
        // flush reference-particle forces (D/Partial_Force_Acc.v:65-129)
        if (F32) {
            for (int k = tid; k < 3 * mi; k += NTHR) {
                float v = s_iacc_f[k];
                if (v != 0.0f) atomic_add_agent(&a.f32[4 * (size_t)(hstart + itile + k / 3) + k % 3], v);
            }
        } else {
            long long *dst = a.fix + 3 * (size_t)(hstart + itile);
            for (int k = tid; k < 3 * mi; k += NTHR) {
                unsigned long long v = s_iacc[k];
                if (v) atomic_add_agent(reinterpret_cast<unsigned long long *>(dst + k), v);
            }
        }
    }
    // per-wave statistics (npairs/ncand are wave-uniform)
    if (lane == 0) {
        atomic_add_agent(&a.counters[0], npairs);
        atomic_add_agent(&a.counters[1], ncand * 64ull);
    }
    if (flag) __hip_atomic_fetch_or(&a.counters[2], (unsigned long long)flag, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
}

// Offset groups: the home cell is half the work of a full neighbour cell.
static void offset_groups(int ngroups, int *gstart)
{
    static const int g1[2] = {0, 14};
    static const int g2[3] = {0, 7, 14};
    static const int g4[5] = {0, 4, 8, 11, 14};
    const int *g = ngroups == 1 ? g1 : ngroups == 2 ? g2 : g4;
    for (int k = 0; k <= ngroups; ++k) gstart[k] = g[k];
}

int launch_force(DeviceSystem &s, hipStream_t st)
{
    ForceArgs a;
    a.pos = s.pos; a.type = s.type; a.cell_start = s.cell_start;
    a.nc0 = s.g.nc[0]; a.nc1 = s.g.nc[1]; a.nc2 = s.g.nc[2];
    a.box0 = s.g.box[0]; a.box1 = s.g.box[1]; a.box2 = s.g.box[2];
    a.rc2 = s.g.rc2;
    a.tab = s.tab; a.off = s.exp_offset; a.nseg = s.nseg;
    a.win_lo = s.win.lo; a.nwin = s.win.n;
    a.pair_ab = s.pair_ab; a.ntypes = s.ntypes;
    a.fix = s.fix; a.f32 = s.f32; a.counters = s.counters;
    a.ngroups = 4;
    offset_groups(a.ngroups, a.gstart);
    const bool f32 = s.flags & MDRL_ACC_FP32, par = s.flags & MDRL_EVAL_PARITY,
               ana = s.flags & MDRL_EVAL_ANALYTIC, coul = s.coul;
    const int str = coul ? 6 : 4;
    size_t lds = ana ? 0 : (size_t)s.win.n * str * sizeof(float);
    dim3 grid(s.g.ncells * a.ngroups), block(NTHR);
#define MDRL_LAUNCH_RL(C, F, P, A)                                                            \
    hipLaunchKernelGGL((k_rl_halfshell<C, F, P, A>), grid, block, lds, st, a)
    if (ana) {
        if (coul) { if (f32) MDRL_LAUNCH_RL(true, true, false, true); else MDRL_LAUNCH_RL(true, false, false, true); }
        else { if (f32) MDRL_LAUNCH_RL(false, true, false, true); else MDRL_LAUNCH_RL(false, false, false, true); }
    } else if (par) {
        if (coul) { if (f32) MDRL_LAUNCH_RL(true, true, true, false); else MDRL_LAUNCH_RL(true, false, true, false); }
        else { if (f32) MDRL_LAUNCH_RL(false, true, true, false); else MDRL_LAUNCH_RL(false, false, true, false); }
    } else {
        if (coul) { if (f32) MDRL_LAUNCH_RL(true, true, false, false); else MDRL_LAUNCH_RL(true, false, false, false); }
        else { if (f32) MDRL_LAUNCH_RL(false, true, false, false); else MDRL_LAUNCH_RL(false, false, false, false); }
    }
#undef MDRL_LAUNCH_RL
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// binning: counting sort by cell, made stable by a per-cell rank on the input index
// (M/Simulation_Verification_LJArgon.m:143-175 keeps atoms in input order per cell).
// ---------------------------------------------------------------------------------

__global__ void k_assign(int n, const float4 *pos, float ie0, float ie1, float ie2, int nc0,
                         int nc1, int nc2, int *cell, int *count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (s < n) {
        const float4 p = pos[s];
        c = (cell_of(p.x, ie0, nc0) * nc1 + cell_of(p.y, ie1, nc1)) * nc2 + cell_of(p.z, ie2, nc2);
        cell[s] = c;
    }
    run_add(count, c, s < n);
}

// Exclusive scan of the next-step counts into cell_start; clears counts and fill
// pointers; records the largest cell.  One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan(int ncells, int *count, int *start, int *fill,
                                               unsigned long long *counters)
{
    __shared__ int s_sum[1024];
    __shared__ int s_max[1024];
    const int t = threadIdx.x;
    const int per = (ncells + 1023) / 1024;
    const int b = t * per, e = min(ncells, b + per);
    int sum = 0, mx = 0;
    for (int c = b; c < e; ++c) { int v = count[c]; sum += v; mx = max(mx, v); }
    s_sum[t] = sum;
    s_max[t] = mx;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? s_sum[t - off] : 0;
        int m = (t >= off) ? s_max[t - off] : 0;
        __syncthreads();
        s_sum[t] += v;
        s_max[t] = max(s_max[t], m);
        __syncthreads();
    }
    int run = s_sum[t] - sum;
    for (int c = b; c < e; ++c) {
        int v = count[c];
        start[c] = run;
        run += v;
        count[c] = 0;
        fill[c] = 0;
    }
    if (t == 1023) {
        start[ncells] = s_sum[1023];
        counters[3] = (unsigned long long)s_max[1023];
        counters[0] = 0ull;   // pair and candidate counts of the force evaluation that follows
        counters[1] = 0ull;
    }
}

__global__ void k_scatter(int n, const int *cell, const int *start, int *fill, int *perm)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (s < n) ? cell[s] : 0;
    const int k = run_add(fill, c, s < n);
    if (s < n) perm[start[c] + k] = s;
}

// Orders each cell's slots by input index.  Cells larger than the LDS capacity keep
// the scatter order and raise MDRL_F_CELL_OVERFLOW (forces stay correct in int64 mode).
constexpr int SORT_CAP = 2048;
__global__ __launch_bounds__(256) void k_cell_sort(const int *start, int *perm, const int *orig,
                                                   unsigned long long *counters)
{
    __shared__ int s_v[SORT_CAP];
    __shared__ int s_k[SORT_CAP];
    const int c = blockIdx.x;
    const int b = start[c], m = start[c + 1] - b;
    if (m <= 1) return;
    if (m > SORT_CAP) {
        if (threadIdx.x == 0)
            __hip_atomic_fetch_or(&counters[2], 4ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        int v = perm[b + k];
        s_v[k] = v;
        s_k[k] = orig[v];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        int key = s_k[k], r = 0;
        for (int u = 0; u < m; ++u) r += (s_k[u] < key);
        perm[b + r] = s_v[k];
    }
}

__global__ void k_gather(int n, const int *perm, const float4 *pos, const float4 *vel,
                         const int *type, const int *orig, float4 *pos2, float4 *vel2,
                         int *type2, int *orig2, long long *fix, float *f32)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int o = perm[s];
    pos2[s] = pos[o];
    vel2[s] = vel[o];
    type2[s] = type[o];
    orig2[s] = orig[o];
    if (fix) { fix[3 * (size_t)s] = 0; fix[3 * (size_t)s + 1] = 0; fix[3 * (size_t)s + 2] = 0; }
    if (f32) reinterpret_cast<float4 *>(f32)[s] = make_float4(0.f, 0.f, 0.f, 0.f);
}

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

int launch_bin_initial(DeviceSystem &s, hipStream_t st)
{
    const Geometry &g = s.g;
    if (g.n > 0)
        hipLaunchKernelGGL(k_assign, dim3(nblk(g.n, 256)), dim3(256), 0, st, g.n, s.pos,
                           g.inv_edge[0], g.inv_edge[1], g.inv_edge[2], g.nc[0], g.nc[1], g.nc[2],
                           s.cell, s.cell_next);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

int launch_rebin(DeviceSystem &s, hipStream_t st)
{
    const Geometry &g = s.g;
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, g.ncells, s.cell_next, s.cell_start,
                       s.cell_fill, s.counters);
    if (g.n > 0) {
        hipLaunchKernelGGL(k_scatter, dim3(nblk(g.n, 256)), dim3(256), 0, st, g.n, s.cell,
                           s.cell_start, s.cell_fill, s.perm);
        hipLaunchKernelGGL(k_cell_sort, dim3(g.ncells), dim3(256), 0, st, s.cell_start, s.perm,
                           s.orig, s.counters);
        hipLaunchKernelGGL(k_gather, dim3(nblk(g.n, 256)), dim3(256), 0, st, g.n, s.perm, s.pos,
                           s.vel, s.type, s.orig, s.pos2, s.vel2, s.type2, s.orig2,
                           (s.flags & MDRL_ACC_FP32) ? nullptr : s.fix,
                           (s.flags & MDRL_ACC_FP32) ? s.f32 : nullptr);
    }
    MDRL_HIP(hipGetLastError());
    std::swap(s.pos, s.pos2);
    std::swap(s.vel, s.vel2);
    std::swap(s.type, s.type2);
    std::swap(s.orig, s.orig2);
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// k_update: fused Motion_Update (D/Motion_Update.v:685-805) as velocity Verlet:
//   dv  = (float)(fix * hk[type])            hk = 2^-32 * 1e-5 / m * dt/2
//   v   = vh + dv                            (full-step velocity; v = vh on the first step)
//   vh' = v + dv;  x' = wrap(x + vh'*dt);  next cell of x' counted for the rebin.
// MDRL_KICK_DRIFT gives the reference's kick-then-drift: v' = v + (dv + dv).
// The kinetic energy of the full-step velocity is reduced in FP64.
// ---------------------------------------------------------------------------------

// Periodic wrap into [0, L): one add or subtract of L (the reference's wrap, C model
// :175-192), then a floor-based wrap for the rare atom that moved more than a box
// length; a non-finite coordinate becomes 0 (the caller has flagged it already).
__device__ __forceinline__ float wrap1(float x, float L)
{
    if (x < 0.0f) x = x + L;
    if (x >= L) x = x - L;
    if (!(x >= 0.0f && x < L)) {
        x = x - L * floorf(x / L);
        if (!(x >= 0.0f && x < L)) x = 0.0f;
    }
    return x;
}

__global__ __launch_bounds__(256) void k_update(int n, float4 *pos, float4 *vel, const int *type,
                                                const long long *fix, const float *f32,
                                                const double *hk, const double *mass, float dt,
                                                float b0, float b1, float b2, float ie0, float ie1,
                                                float ie2, int nc0, int nc1, int nc2, int first,
                                                int kick_drift, int *cell, int *count,
                                                double *red, unsigned long long *counters)
{
    __shared__ double s_red[256];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    unsigned flag = 0;
    int c = 0;
    if (s < n) {
        const int t = type[s];
        const double k = hk[t];
        float f3[3];
        if (f32) {
            float4 f = reinterpret_cast<const float4 *>(f32)[s];
            f3[0] = (float)((double)f.x * 4294967296.0 * k);
            f3[1] = (float)((double)f.y * 4294967296.0 * k);
            f3[2] = (float)((double)f.z * 4294967296.0 * k);
        } else {
            f3[0] = (float)((double)fix[3 * (size_t)s + 0] * k);
            f3[1] = (float)((double)fix[3 * (size_t)s + 1] * k);
            f3[2] = (float)((double)fix[3 * (size_t)s + 2] * k);
        }
        float4 p = pos[s], v = vel[s];
        float vi[3] = {v.x, v.y, v.z}, xi[3] = {p.x, p.y, p.z};
        const float L[3] = {b0, b1, b2};
        double vv = 0.0;
        for (int d = 0; d < 3; ++d) {
            float dv = f3[d];
            float vn;
            if (kick_drift) {
                float dv2 = dv + dv;
                vn = vi[d] + dv2;
                vv += (double)vn * (double)vn;
            } else {
                float vf = first ? vi[d] : vi[d] + dv;
                vv += (double)vf * (double)vf;
                vn = vf + dv;
            }
            float step = vn * dt;
            float x = xi[d] + step;
            if (!isfinite(x) || !isfinite(vn)) flag |= 2u;
            xi[d] = wrap1(x, L[d]);
            vi[d] = vn;
        }
        ke = 0.5 * mass[t] * vv * 1.0e4;
        pos[s] = make_float4(xi[0], xi[1], xi[2], p.w);
        vel[s] = make_float4(vi[0], vi[1], vi[2], (float)ke);
        c = (cell_of(xi[0], ie0, nc0) * nc1 + cell_of(xi[1], ie1, nc1)) * nc2 + cell_of(xi[2], ie2, nc2);
        cell[s] = c;
    }
    run_add(count, c, s < n);
    s_red[threadIdx.x] = ke;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_add_agent(&red[0], s_red[0]);
    if (flag) __hip_atomic_fetch_or(&counters[2], (unsigned long long)flag, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
}

int launch_update(DeviceSystem &s, float dt, hipStream_t st)
{
    const Geometry &g = s.g;
    if (g.n == 0) return MDRL_OK;
    const bool f32 = s.flags & MDRL_ACC_FP32;
    hipLaunchKernelGGL(k_update, dim3(nblk(g.n, 256)), dim3(256), 0, st, g.n, s.pos, s.vel, s.type,
                       f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr, s.hk, s.mass, dt, g.box[0],
                       g.box[1], g.box[2], g.inv_edge[0], g.inv_edge[1], g.inv_edge[2], g.nc[0],
                       g.nc[1], g.nc[2], s.first_step ? 1 : 0,
                       (s.flags & MDRL_KICK_DRIFT) ? 1 : 0, s.cell, s.cell_next, s.red, s.counters);
    MDRL_HIP(hipGetLastError());
    s.first_step = false;
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// k_minimize: capped steepest descent, the untimed pre-relaxation of lattice systems
// whose jitter leaves overlapping atoms.  Per atom: F in kJ/mol/nm (int64 * 2^-32),
// s = min(gamma, dmax / |F|) in double, x' = wrap(x + (float)(F * s)).
// ---------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_minimize(int n, float4 *pos, const long long *fix,
                                                  const float *f32, double gamma, double dmax,
                                                  float b0, float b1, float b2, float ie0,
                                                  float ie1, float ie2, int nc0, int nc1, int nc2,
                                                  int *cell, int *count,
                                                  unsigned long long *counters)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned flag = 0;
    int c = 0;
    if (s < n) {
        double f[3];
        for (int d = 0; d < 3; ++d)
            f[d] = f32 ? (double)f32[4 * (size_t)s + d]
                       : (double)fix[3 * (size_t)s + d] * (1.0 / 4294967296.0);
        const double fm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        const double sc = (fm * gamma > dmax) ? dmax / fm : gamma;
        float4 p = pos[s];
        float xi[3] = {p.x, p.y, p.z};
        const float L[3] = {b0, b1, b2};
        for (int d = 0; d < 3; ++d) {
            float step = (float)(f[d] * sc);
            float x = xi[d] + step;
            if (!isfinite(x)) flag |= 2u;
            xi[d] = wrap1(x, L[d]);
        }
        pos[s] = make_float4(xi[0], xi[1], xi[2], p.w);
        c = (cell_of(xi[0], ie0, nc0) * nc1 + cell_of(xi[1], ie1, nc1)) * nc2 + cell_of(xi[2], ie2, nc2);
        cell[s] = c;
    }
    run_add(count, c, s < n);
    if (flag) __hip_atomic_fetch_or(&counters[2], (unsigned long long)flag, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
}

int launch_minimize(DeviceSystem &s, double gamma, double dmax, hipStream_t st)
{
    const Geometry &g = s.g;
    if (g.n == 0) return MDRL_OK;
    const bool f32 = s.flags & MDRL_ACC_FP32;
    hipLaunchKernelGGL(k_minimize, dim3(nblk(g.n, 256)), dim3(256), 0, st, g.n, s.pos,
                       f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr, gamma, dmax, g.box[0],
                       g.box[1], g.box[2], g.inv_edge[0], g.inv_edge[1], g.inv_edge[2], g.nc[0],
                       g.nc[1], g.nc[2], s.cell, s.cell_next, s.counters);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// k_energy: potential energy and virial from the c*_{12,6,1} tables
// (M/FullScale_Simulation_LJArgon.m:938-1010), half shell, FP64 reduction.  One
// workgroup per home cell, one home atom per thread, no compaction: it runs once per
// energy report, not per step.
// ---------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_energy(ForceArgs a, const float *etab, const float2 *ab_e,
                                                double *red)
{
    __shared__ double s_r[10][256];
    const int h = blockIdx.x;
    const int hz = h % a.nc2, hy = (h / a.nc2) % a.nc1, hx = h / (a.nc1 * a.nc2);
    const int hstart = a.cell_start[h], hcount = a.cell_start[h + 1] - hstart;
    double e = 0.0, w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int ia = threadIdx.x; ia < hcount; ia += blockDim.x) {
        const float4 pi = a.pos[hstart + ia];
        const int ti = a.type[hstart + ia];
        for (int o = 0; o < 14; ++o) {
            int rx = hx + c_hs[o][0], ry = hy + c_hs[o][1], rz = hz + c_hs[o][2];
            float sx = 0.f, sy = 0.f, sz = 0.f;
            if (rx < 0) { rx += a.nc0; sx = -a.box0; } else if (rx >= a.nc0) { rx -= a.nc0; sx = a.box0; }
            if (ry < 0) { ry += a.nc1; sy = -a.box1; } else if (ry >= a.nc1) { ry -= a.nc1; sy = a.box1; }
            if (rz < 0) { rz += a.nc2; sz = -a.box2; } else if (rz >= a.nc2) { rz -= a.nc2; sz = a.box2; }
            const int c = (rx * a.nc1 + ry) * a.nc2 + rz;
            const int cs = a.cell_start[c], cc = a.cell_start[c + 1] - cs;
            for (int jb = (o == 0 ? ia + 1 : 0); jb < cc; ++jb) {
                float4 pj = a.pos[cs + jb];
                pj.x = pj.x + sx; pj.y = pj.y + sy; pj.z = pj.z + sz;
                float dx, dy, dz;
                float r2 = pair_geom<false>(pi, pj, 0.f, 0.f, 0.f, dx, dy, dz);
                if (!(r2 > 0.0f && r2 <= a.rc2)) continue;
                const int tj = a.type[cs + jb];
                // ab_e = (eps sig^12, eps sig^6): E = A t12 - B t6, F = A t14 - B t8 (unit
                // tables 4 r^-12, 4 r^-6, 48 r^-14, 24 r^-8); the scaled charges carry 10x.
                const float2 abe = ab_e[ti * a.ntypes + tj];
                const int idx = lut_seg(r2, a.off, a.nseg) * 256 + lut_bin(r2);
                const float *pe = etab + (size_t)idx * 6;
                const float *pf = a.tab + (size_t)idx * 6;
                double t12 = (double)pe[1] * r2 + pe[0], t6 = (double)pe[3] * r2 + pe[2];
                double E = abe.x * t12 - abe.y * t6;
                double t14 = (double)pf[1] * r2 + pf[0], t8 = (double)pf[3] * r2 + pf[2];
                double F = abe.x * t14 - abe.y * t8;
                if (pi.w != 0.0f && pj.w != 0.0f) {
                    double qq = (double)pi.w * pj.w;
                    E += qq * ((double)pe[5] * r2 + pe[4]) * 0.1;
                    F += qq * ((double)pf[5] * r2 + pf[4]) * 0.1;
                }
                e += E;
                const double d3[3] = {dx, dy, dz};
                for (int u = 0; u < 3; ++u)
                    for (int v = 0; v < 3; ++v) w[3 * u + v] += F * d3[u] * d3[v];
            }
        }
    }
    s_r[0][threadIdx.x] = e;
    for (int k = 0; k < 9; ++k) s_r[k + 1][threadIdx.x] = w[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 10; ++k) s_r[k][threadIdx.x] += s_r[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 10) atomic_add_agent(&red[1 + threadIdx.x], s_r[threadIdx.x][0]);
}

int launch_energy(DeviceSystem &s, hipStream_t st)
{
    ForceArgs a;
    a.pos = s.pos; a.type = s.type; a.cell_start = s.cell_start;
    a.nc0 = s.g.nc[0]; a.nc1 = s.g.nc[1]; a.nc2 = s.g.nc[2];
    a.box0 = s.g.box[0]; a.box1 = s.g.box[1]; a.box2 = s.g.box[2];
    a.rc2 = s.g.rc2;
    a.tab = s.tab; a.off = s.exp_offset; a.nseg = s.nseg;
    a.win_lo = s.win.lo; a.nwin = s.win.n;
    a.pair_ab = s.pair_ab; a.ntypes = s.ntypes;
    a.fix = s.fix; a.f32 = s.f32; a.counters = s.counters; a.ngroups = 1;
    hipLaunchKernelGGL(k_energy, dim3(s.g.ncells), dim3(256), 0, st, a, s.etab, s.pair_ab_e, s.red);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// small state kernels (fills instead of hipMemsetAsync so a captured step stays valid)
// ---------------------------------------------------------------------------------

__global__ void k_zero_forces(int n, long long *fix, float *f32)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    if (fix) { fix[3 * (size_t)s] = 0; fix[3 * (size_t)s + 1] = 0; fix[3 * (size_t)s + 2] = 0; }
    if (f32) reinterpret_cast<float4 *>(f32)[s] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0ull;
}

__global__ void k_zero_f64(double *p, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

// Full-step velocity v = vh + dv(F) (or vh before the first step) into vout (input order
// is restored by the host) and its FP64 kinetic energy into red[0].
__global__ __launch_bounds__(256) void k_kinetic(int n, const float4 *vel, const int *type,
                                                 const long long *fix, const float *f32,
                                                 const double *hk, const double *mass,
                                                 int first, float4 *vout, double *red)
{
    __shared__ double s_red[256];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    if (s < n) {
        const int t = type[s];
        const double k = hk[t];
        float4 v = vel[s];
        float vi[3] = {v.x, v.y, v.z};
        double vv = 0.0;
        for (int d = 0; d < 3; ++d) {
            float dv = f32 ? (float)((double)f32[4 * (size_t)s + d] * 4294967296.0 * k)
                           : (float)((double)fix[3 * (size_t)s + d] * k);
            float vf = first ? vi[d] : vi[d] + dv;
            vi[d] = vf;
            vv += (double)vf * (double)vf;
        }
        ke = 0.5 * mass[t] * vv * 1.0e4;
        vout[s] = make_float4(vi[0], vi[1], vi[2], (float)ke);
    }
    s_red[threadIdx.x] = ke;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_add_agent(&red[0], s_red[0]);
}

int launch_zero_forces(DeviceSystem &s, hipStream_t st, bool forces)
{
    if (!forces) return MDRL_OK;   // after a rebin: forces and pair counters are already zero
    if (s.g.n > 0) {
        const bool f32 = s.flags & MDRL_ACC_FP32;
        hipLaunchKernelGGL(k_zero_forces, dim3(nblk(s.g.n, 256)), dim3(256), 0, st, s.g.n,
                           f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr);
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, s.counters, 2);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

int launch_zero_red(DeviceSystem &s, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, s.red, 11);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

int launch_kinetic(DeviceSystem &s, float4 *vout, hipStream_t st)
{
    if (s.g.n == 0) return MDRL_OK;
    const bool f32 = s.flags & MDRL_ACC_FP32;
    hipLaunchKernelGGL(k_kinetic, dim3(nblk(s.g.n, 256)), dim3(256), 0, st, s.g.n, s.vel, s.type,
                       f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr, s.hk, s.mass,
                       s.first_step ? 1 : 0, vout, s.red);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// k_eval_pairs: the pair-batch API (O/test_hdl_lib/device/LJ.cl:31-80 calling
// RL_LJ_Evaluation, O/HDL_Source/RL_LJ_Evaluation_OpenCL_Top.v).  One pair per thread,
// table read through the cache hierarchy.
// ---------------------------------------------------------------------------------

__global__ void k_eval_pairs(const float *tab, int off, int nseg, const float *ref,
                             const float *nbr, const float *pa, const float *pb,
                             const float *pqq, float *fs, float *fxyz, long long n, float rc2,
                             int flags)
{
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const bool par = flags & MDRL_PAIR_PARITY, strict = flags & MDRL_PAIR_STRICT,
               coul = flags & MDRL_PAIR_COULOMB;
    float dx = ref[3 * p] - nbr[3 * p];
    float dy = ref[3 * p + 1] - nbr[3 * p + 1];
    float dz = ref[3 * p + 2] - nbr[3 * p + 2];
    float r2;
    if (par) { r2 = dx * dx + dy * dy; r2 = r2 + dz * dz; }
    else r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    bool in = (r2 > 0.0f) && (strict ? (r2 < rc2) : (r2 <= rc2));
    float F = 0.f, fx = 0.f, fy = 0.f, fz = 0.f;
    if (in) {
        const int idx = lut_seg(r2, off, nseg) * 256 + lut_bin(r2);
        const float *c = tab + (size_t)idx * 6;
        const float A = pa ? pa[p] : 1.0f, B = pb ? pb[p] : 1.0f, Q = pqq ? pqq[p] : 0.0f;
        if (par) {
            float t14 = c[1] * r2; t14 = t14 + c[0];
            float t8 = c[3] * r2;  t8 = t8 + c[2];
            if (pa || pb) { float x = A * t14, y = B * t8; F = x - y; }
            else F = t14 - t8;
            if (coul) { float t3 = c[5] * r2; t3 = t3 + c[4]; float z = Q * t3; F = F + z; }
        } else {
            float t14 = fmaf(c[1], r2, c[0]);
            float t8 = fmaf(c[3], r2, c[2]);
            F = fmaf(A, t14, -(B * t8));
            if (coul) { float t3 = fmaf(c[5], r2, c[4]); F = fmaf(Q, t3, F); }
        }
        fx = F * dx; fy = F * dy; fz = F * dz;
    }
    if (fs) fs[p] = F;
    fxyz[3 * p] = fx; fxyz[3 * p + 1] = fy; fxyz[3 * p + 2] = fz;
}

int launch_eval_pairs(const float *d_tab, int exp_offset, int nseg, const float *ref,
                      const float *nbr, const float *pa, const float *pb, const float *pqq,
                      float *fs, float *fxyz, int64_t n, float rc2, int flags, hipStream_t st)
{
    if (n <= 0) return MDRL_OK;
    long long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_eval_pairs, dim3((unsigned)blocks), dim3(256), 0, st, d_tab, exp_offset,
                       nseg, ref, nbr, pa, pb, pqq, fs, fxyz, (long long)n, rc2, flags);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// The C model on the GPU: R/C_Model_ForceEval/MD_Evaluation_Model.cpp:57-203 with the
// same promotions (double pow, double +-0.5 in the PBC, double 0.5*MASS), one thread
// per reference particle, neighbours summed in index order as the reference does.
// ---------------------------------------------------------------------------------

struct CMConsts {
    float s12, s6, c4, c48, c24, rc2, mass, dt, box0, box1, box2;
};

__device__ __forceinline__ float cm_pbc(float d, float L)
{
    float q = d / L;
    int k = (d >= 0) ? (int)((double)q + 0.5) : (int)((double)q - 0.5);
    return d - L * (float)k;
}

__global__ __launch_bounds__(64) void k_cm_force(int n, const float *x, float *v, float *f, float *e,
                                                float *ke, CMConsts c)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float Evdw_acc = 0, Fx = 0, Fy = 0, Fz = 0;
    const float rx = x[r], ry = x[n + r], rz = x[2 * n + r];
    for (int nb = 0; nb < n; ++nb) {
        float dx = rx - x[nb], dy = ry - x[n + nb], dz = rz - x[2 * n + nb];
        dx = cm_pbc(dx, c.box0);
        dy = cm_pbc(dy, c.box1);
        dz = cm_pbc(dz, c.box2);
        float r2 = dx * dx + dy * dy;
        r2 = r2 + dz * dz;
        if (r2 > 0 && r2 <= c.rc2) {
            float inv_r2 = 1.0f / r2;
            float inv_r8 = (float)pow((double)inv_r2, 4.0);
            float inv_r14 = (float)pow((double)inv_r2, 7.0);
            float inv_r6 = (float)pow((double)inv_r2, 3.0);
            float inv_r12 = (float)pow((double)inv_r2, 6.0);
            float vdw12 = c.c4 * c.s12 * inv_r12;
            float vdw6 = c.c4 * c.s6 * inv_r6;
            float vdw14 = c.c48 * c.s12 * inv_r14;
            float vdw8 = c.c24 * c.s6 * inv_r8;
            float Fvdw = vdw14 - vdw8;
            float Evdw = vdw12 - vdw6;
            float t;
            t = Fvdw * dx; Fx += t;
            t = Fvdw * dy; Fy += t;
            t = Fvdw * dz; Fz += t;
            Evdw_acc += Evdw;
        }
    }
    e[r] = Evdw_acc;
    f[r] = Fx; f[n + r] = Fy; f[2 * n + r] = Fz;
    float ax = Fx / c.mass, ay = Fy / c.mass, az = Fz / c.mass;
    float vx = v[r], vy = v[n + r], vz = v[2 * n + r];
    float t;
    t = ax * c.dt; vx += t;
    t = ay * c.dt; vy += t;
    t = az * c.dt; vz += t;
    float v2 = vx * vx + vy * vy;
    v2 = v2 + vz * vz;
    ke[r] = (float)((0.5 * (double)c.mass) * (double)v2);
    v[r] = vx; v[n + r] = vy; v[2 * n + r] = vz;
}

// System_Energy / System_Kinetic_Energy accumulated in reference order (:154-158).
__global__ void k_cm_sum(int n, const float *e, const float *ke, float *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float se = 0, sk = 0;
    for (int r = 0; r < n; ++r) { se += e[r]; sk += ke[r]; }
    out[0] = (float)((double)se * 0.5);
    out[1] = sk;
}

__global__ void k_cm_drift(int n, float *x, const float *v, float *f, CMConsts c)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float L[3] = {c.box0, c.box1, c.box2};
    for (int d = 0; d < 3; ++d) {
        float p = x[d * n + i];
        float t = v[d * n + i] * c.dt;
        p += t;
        if (p < 0) p += L[d]; else if (p >= L[d]) p -= L[d];
        x[d * n + i] = p;
        f[d * n + i] = 0;
    }
}

int launch_cmodel_iteration(CModelDev &m, const md_rl_cmodel_cfg &cfg, int it, hipStream_t st)
{
    CMConsts c;
    c.s12 = (float)pow((double)cfg.sigma, 12.0);
    c.s6 = (float)pow((double)cfg.sigma, 6.0);
    c.c4 = 4.0f * cfg.eps;
    c.c48 = 48.0f * cfg.eps;
    c.c24 = 24.0f * cfg.eps;
    c.rc2 = cfg.rc * cfg.rc;
    c.mass = cfg.mass;
    c.dt = cfg.dt;
    c.box0 = cfg.box[0]; c.box1 = cfg.box[1]; c.box2 = cfg.box[2];
    hipLaunchKernelGGL(k_cm_force, dim3(nblk(m.n, 64)), dim3(64), 0, st, m.n, m.x, m.v, m.f, m.e,
                       m.ke, c);
    hipLaunchKernelGGL(k_cm_sum, dim3(1), dim3(64), 0, st, m.n, m.e, m.ke, m.sums + 2 * it);
    hipLaunchKernelGGL(k_cm_drift, dim3(nblk(m.n, 256)), dim3(256), 0, st, m.n, m.x, m.v, m.f, c);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

}  // namespace mdrl
