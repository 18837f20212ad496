// mdrl_system.hip -- gfx950 kernels of the half-shell system path: force evaluation,
// binning (counting sort), Motion_Update, pre-relaxation and FP64 energy reductions, for a
// batch of independent periodic systems (a single system is a batch of one).
//
// Reference paths: D/ = Range_Limited_demo/SourceCode/, R/ = Range_Limited/,
// M/ = D/MatlabScripts/, O/ = D/OpenCL_HDL_Lib/.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "mdrl_device.h"

namespace mdrl {

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

// ---------------------------------------------------------------------------------
// Bonded terms (Bonded/Bonded_Top.v:202-380 streams bonds, angles and torsions through
// calc_bond_v2.v, calc_angle_v2.v and calcTorsion_v2.v): one thread per term, FP32 forces
// in kJ/mol/nm (the constants carry the factor 10), added into the same fixed-point
// buffer as the range-limited forces.  Terms name atoms by input index; slot_of maps
// them to the current slots.  The arithmetic is written operation by operation (the
// build does not contract), so the oracle restates it bit for bit.
// ---------------------------------------------------------------------------------

struct BondedArgs {
    const float4 *pos;
    const int *slot_of;
    const int *slot_sys;
    const SysGeom *geom;
    const int4 *atoms;
    const float4 *params;
    int nterms;
    long long *fix;
    float *f32;
    double *term_e;
    unsigned long long *counters;
    const int4 *blocks;       // per block: first term, end term, first atom, atom span
    int nblocks;
};

__device__ __forceinline__ float img_f(float d, float L)
{
    const float h = 0.5f * L;
    if (d > h) return d - L;
    if (d < -h) return d + L;
    return d;
}
__device__ __forceinline__ float dot3f(const float *a, const float *b)
{
    float r = a[0] * b[0] + a[1] * b[1];
    return r + a[2] * b[2];
}
__device__ __forceinline__ void cross3f(const float *a, const float *b, float *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// atan2 for the bond angle: reduce to a = min/max in [0, 1]; above tan(pi/12) use
// atan a = pi/6 + atan((a sqrt3 - 1)/(a + sqrt3)); then the Taylor series of atan to
// t^11 (|t| <= 0.268, next term < 3e-9).
__device__ __forceinline__ float angle_atan2f(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    const float a = mn / mx;
    float t = a, off = 0.0f;
    if (a > 0.26794919f) {
        const float num = a * 1.7320508f - 1.0f;
        const float den = a + 1.7320508f;
        t = num / den;
        off = 0.52359878f;
    }
    const float t2 = t * t;
    float p = -0.09090909f;
    p = p * t2 + 0.11111111f;
    p = p * t2 - 0.14285715f;
    p = p * t2 + 0.2f;
    p = p * t2 - 0.33333334f;
    p = p * t2 + 1.0f;
    float r = p * t;
    r = off + r;
    if (ay > ax) r = 1.5707964f - r;
    if (x < 0.0f) r = 3.1415927f - r;
    return y < 0.0f ? -r : r;
}

// The forces of a term's atoms, added atomically.  In fixed point the last atom takes
// minus the sum of the others' converted forces, so each term's forces sum to zero exactly.
// Forces go to a sink: the LDS accumulators of the term block (atom index - first atom of
// the block), flushed to HBM once per atom by k_bonded or by the bonded tasks of
// k_rl_halfshell.  The terms of a block touch at most kBondedBlockSpan consecutive input
// indices.
struct BondedSink {
    unsigned long long *acc;   // 3 per atom of the block
    float *accf;               // FP32-atomics mode: 3 per atom
    int lo;                    // first input index of the block
};

template <int NA>
__device__ __forceinline__ void bonded_add(const BondedSink &k_, const int *atom, const float (*f)[3],
                                           unsigned &flag)
{
    if (k_.accf) {
        for (int k = 0; k < NA; ++k)
            for (int d = 0; d < 3; ++d) atomic_add_wg(&k_.accf[3 * (atom[k] - k_.lo) + d], f[k][d]);
        return;
    }
    unsigned long long sum[3] = {0ull, 0ull, 0ull};
    for (int k = 0; k < NA; ++k) {
        long long q[3];
        if (k < NA - 1) {
            fix3(f[k][0], f[k][1], f[k][2], q[0], q[1], q[2], flag);
            for (int d = 0; d < 3; ++d) sum[d] += (unsigned long long)q[d];
        } else {
            for (int d = 0; d < 3; ++d) q[d] = (long long)(0ull - sum[d]);
        }
        for (int d = 0; d < 3; ++d) atomic_add_wg(&k_.acc[3 * (atom[k] - k_.lo) + d], (unsigned long long)q[d]);
    }
}

// One term.  ENERGY: return the FP64 energy of the term (from the same FP32
// intermediates) instead of adding forces.
template <bool ENERGY>
__device__ __forceinline__ double bonded_term(const BondedArgs &a, int t, const BondedSink &sink,
                                              unsigned &flag)
{
    const int4 at = a.atoms[t];
    const float4 pp = a.params[t];
    const int s0 = a.slot_of[at.x], s1 = a.slot_of[at.y];
    const SysGeom &G = a.geom[a.slot_sys[s0]];
    const float4 p1 = a.pos[s0], p2 = a.pos[s1];
    const float x1[3] = {p1.x, p1.y, p1.z}, x2[3] = {p2.x, p2.y, p2.z};
    double e = 0.0;
    if (at.z < 0) {
        // bond, Bonded/calc_bond_v2.v:83-110
        float r12[3];
        for (int d = 0; d < 3; ++d) r12[d] = img_f(x1[d] - x2[d], G.box[d]);
        const float r = sqrtf(dot3f(r12, r12));
        const float dr = r - pp.x;
        if (ENERGY) {
            const double ddr = (double)r - (double)pp.x;
            e = (double)pp.y / 10.0 * ddr * ddr;
        } else {
            float dp = 2.0f * pp.y;
            dp = dp * dr;
            const float c = -dp / r;
            float f[2][3];
            for (int d = 0; d < 3; ++d) { f[0][d] = c * r12[d]; f[1][d] = -f[0][d]; }
            const int ids[2] = {at.x, at.y};
            bonded_add<2>(sink, ids, f, flag);
        }
    } else if (at.w < 0) {
        // angle, Bonded/calc_angle_v2.v:113-147 (no Urey-Bradley term)
        const int s2 = a.slot_of[at.z];
        const float4 p3 = a.pos[s2];
        const float x3[3] = {p3.x, p3.y, p3.z};
        float r12[3], r32[3], cr[3];
        for (int d = 0; d < 3; ++d) {
            r12[d] = img_f(x1[d] - x2[d], G.box[d]);
            r32[d] = img_f(x3[d] - x2[d], G.box[d]);
        }
        const float d12 = sqrtf(dot3f(r12, r12)), d32 = sqrtf(dot3f(r32, r32));
        cross3f(r12, r32, cr);
        const float y = sqrtf(dot3f(cr, cr)), xx = dot3f(r12, r32);
        const float th = angle_atan2f(y, xx);
        if (ENERGY) {
            const double dth = (double)th - (double)pp.x;
            e = (double)pp.y / 10.0 * dth * dth;
        } else {
            const float dd = d12 * d32;
            const float cs = xx / dd, sn = y / dd;
            if (sn > 1e-6f) {
                float dp = 2.0f * pp.y;
                dp = dp * (th - pp.x);
                const float c = -dp;
                const float a1 = sn * d12, a3 = sn * d32;
                // order: end 1, end 3, then the vertex (minus their sum)
                float f[3][3];
                for (int d = 0; d < 3; ++d) {
                    const float u12 = r12[d] / d12, u32 = r32[d] / d32;
                    float g1 = u12 * cs - u32;
                    g1 = g1 / a1;
                    float g3 = u32 * cs - u12;
                    g3 = g3 / a3;
                    f[0][d] = c * g1;
                    f[1][d] = c * g3;
                    f[2][d] = -f[0][d] - f[1][d];
                }
                const int ids[3] = {at.x, at.z, at.y};
                bonded_add<3>(sink, ids, f, flag);
            }
        }
    } else {
        // torsion, Bonded/calcTorsion_v2.v:70-274: k (1 + cos(n phi + d)) with the
        // projection-form forces (no division by sin phi)
        const int s2 = a.slot_of[at.z], s3 = a.slot_of[at.w];
        const float4 p3 = a.pos[s2], p4 = a.pos[s3];
        const float x3[3] = {p3.x, p3.y, p3.z}, x4[3] = {p4.x, p4.y, p4.z};
        float rij[3], rkj[3], rkl[3], m[3], v[3];
        for (int d = 0; d < 3; ++d) {
            rij[d] = img_f(x1[d] - x2[d], G.box[d]);
            rkj[d] = img_f(x3[d] - x2[d], G.box[d]);
            rkl[d] = img_f(x3[d] - x4[d], G.box[d]);
        }
        cross3f(rij, rkj, m);
        cross3f(rkj, rkl, v);
        const float m2 = dot3f(m, m), v2 = dot3f(v, v);
        if (m2 > 1e-12f && v2 > 1e-12f) {
            const float kj2 = dot3f(rkj, rkj);
            const float kj = sqrtf(kj2);
            const float mv = sqrtf(m2) * sqrtf(v2);
            const float cp = dot3f(m, v) / mv;
            float sp = kj * dot3f(rij, v);
            sp = sp / mv;
            const int nper = (int)pp.y;
            float cn = cp, snn = sp;
            for (int k = 1; k < nper; ++k) {
                const float c2 = cn * cp - snn * sp;
                const float s2n = snn * cp + cn * sp;
                cn = c2;
                snn = s2n;
            }
            const float sarg = snn * pp.z + cn * pp.w;
            const float carg = cn * pp.z - snn * pp.w;
            if (ENERGY) {
                e = (double)pp.x / 10.0 * (1.0 + (double)carg);
            } else {
                float ddphi = -pp.x * pp.y;
                ddphi = ddphi * sarg;
                float ca = -ddphi * kj;
                ca = ca / m2;
                float cb = ddphi * kj;
                cb = cb / v2;
                const float pq = dot3f(rij, rkj) / kj2, qq = dot3f(rkl, rkj) / kj2;
                // order: i, j, l, then k (minus their sum)
                float f[4][3];
                for (int d = 0; d < 3; ++d) {
                    const float fi = ca * m[d], fl = cb * v[d];
                    const float sv = pq * fi - qq * fl;
                    f[0][d] = fi;
                    f[1][d] = -(fi - sv);
                    f[2][d] = fl;
                    f[3][d] = -(fl + sv);
                }
                const int ids[4] = {at.x, at.y, at.w, at.z};
                bonded_add<4>(sink, ids, f, flag);
            }
        }
    }
    return e;
}

// One block of bonded terms (first term, end term, first atom, atom span), evaluated by
// the threads t0, t0 + nthr, ... of a workgroup into the LDS accumulators acc (3 words per
// atom of the span: fixed point, or FP32 in FP32-atomics mode) and then added into the
// force buffer with one atomic per touched atom and component.  barrier() is the
// workgroup barrier between the phases.
template <typename Barrier>
__device__ __forceinline__ void bonded_block(const BondedArgs &a, int4 blk, unsigned long long *acc, bool f32,
                                             long long *fix, float *fbuf, int t0, int nthr, unsigned &flag,
                                             Barrier barrier)
{
    float *accf = reinterpret_cast<float *>(acc);
    for (int k = t0; k < 3 * kBondedBlockSpan; k += nthr) acc[k] = 0ull;
    barrier();
    const BondedSink sink{f32 ? nullptr : acc, f32 ? accf : nullptr, blk.z};
    for (int t = blk.x + t0; t < blk.y; t += nthr) bonded_term<false>(a, t, sink, flag);
    barrier();
    for (int k = t0; k < 3 * blk.w; k += nthr) {
        const int slot = a.slot_of[blk.z + k / 3];
        if (f32) {
            const float v = accf[k];
            if (v != 0.0f) atomic_add_agent(&fbuf[4 * (size_t)slot + k % 3], v);
        } else {
            const unsigned long long v = acc[k];
            if (v) atomic_add_agent(reinterpret_cast<unsigned long long *>(fix + 3 * (size_t)slot + k % 3), v);
        }
    }
}

// Minimum over the wave: swizzles within each half (no address registers), then the two
// halves.  Exact in any order.
__device__ __forceinline__ float wave_min(float v)
{
    v = fminf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (1 << 10) | 0x1f)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (2 << 10) | 0x1f)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (4 << 10) | 0x1f)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (8 << 10) | 0x1f)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (16 << 10) | 0x1f)));
    return fminf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32)));
}

// Maximum over the wave of an int, the same way.
__device__ __forceinline__ int wave_max(int v)
{
    v = max(v, __builtin_amdgcn_ds_swizzle(v, (1 << 10) | 0x1f));
    v = max(v, __builtin_amdgcn_ds_swizzle(v, (2 << 10) | 0x1f));
    v = max(v, __builtin_amdgcn_ds_swizzle(v, (4 << 10) | 0x1f));
    v = max(v, __builtin_amdgcn_ds_swizzle(v, (8 << 10) | 0x1f));
    v = max(v, __builtin_amdgcn_ds_swizzle(v, (16 << 10) | 0x1f));
    return max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 32));
}

// Inclusive prefix sum over the wave: within each row of 16 lanes by DPP row shifts (lanes
// shifted in from outside the row contribute 0), then the totals of the rows before.
__device__ __forceinline__ int wave_prefix(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    const int r0 = __builtin_amdgcn_readlane(x, 15), r1 = __builtin_amdgcn_readlane(x, 31),
              r2 = __builtin_amdgcn_readlane(x, 47);
    const int row = (int)(threadIdx.x & 63) >> 4;
    return x + (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
}

// ---------------------------------------------------------------------------------
// k_rl_halfshell: one workgroup of 12 waves per task = (home cell, group of half-shell
// offsets), or, in persistent mode, a fixed grid of workgroups pulling tasks (largest
// first) from a counter.  With exclusion windows (bonded systems) the blocks of bonded
// terms are tasks of the same launch, after the range-limited ones, so the two force
// kinds share one persistent grid and the short bonded blocks fill its tail
// (Bonded/Bonded_Top.v:202-380 runs beside the pair pipelines;
// Summation_Logic/Force_Summation_Unit.v:62-126 adds the results).
//
// The home cell's atoms are staged in LDS (tile of TI = 256) and each neighbour cell of
// the group (pre-shifted by its periodic image) as a tile of TJ.  A lane holds one home
// atom in each of up to four chunks of APW atoms (APW = atoms per wavefront: with APW < 64,
// 64/APW lanes share a home atom and split its neighbours).  At step s, lane l tests the
// neighbour (l mod APW) + (l / APW) + s*64/APW (mod m) against all its home atoms, so within
// a step both indices of the queued pairs are distinct and one LDS read feeds every chunk;
// steps are strided over the 12 waves.  The r2 <= rc2 test of D/Filter_Logic.v:234 is a
// ballot; passing pairs are compacted into a per-wave LDS queue (the filter buffers and
// arbiter of D/Filter_Bank.v:548-606) and, after each step, all 64 lanes evaluate each full
// batch of 64 densely: table force from the global table, then the reference particle gets +F and the
// neighbour -F (D/RL_LJ_Evaluation_Unit.v:137) in LDS accumulators, int64 fixed point
// (exact, order-free) or FP32.  Neighbour forces are flushed to HBM with one atomic per
// touched particle and component after each neighbour tile
// (D/Force_Write_Back_Controller.v:370-429), home forces after the group's last offset
// (D/Partial_Force_Acc.v:65-129).
// ---------------------------------------------------------------------------------

// 12 waves per workgroup: two workgroups (24 waves) fill a CU's LDS and its VGPRs at
// 6 waves per SIMD.
constexpr int TI = 256, NWAVE = 12, NTHR = 64 * NWAVE, NCHUNK = 4;
// per-wave pair queue: fewer than 64 entries between steps, up to 64 more per chunk of a step
constexpr int QCAP = 64 + 64 * NCHUNK;
// queue entry: home index (8 bits) | neighbour index (10 bits) << 8 | type pair << 18
constexpr unsigned kJbShift = 8, kJbMask = 0x3ffu, kTpShift = 18;
static_assert(TI <= 256, "queue entries hold 8-bit home indices");

// One neighbour tile: offsets [o, oe) of the task's group, each a whole cell (cells that
// fit the tile are packed together), or a TJ-atom piece starting at atom jt of cell o
// (the home cell, a cell larger than the tile, or any cell in parity mode); mj atoms.
struct NbrTile {
    int o, oe, jt, mj;
};

struct ForceArgs {
    const float4 *pos;
    const int *type;
    const int *cell_start;
    const SysGeom *geom;
    const int *cell_sys;
    const int *tasks;         // task order (persistent mode); NULL: task = blockIdx
    int ntasks;
    float rc2;                // the reference's FP32 cutoff test r2 <= rc2
    float rc2_scan;           // the candidate scan's test on its own r2 (rc2 plus its error bound)
    float rc2_lo;             // analytic mode: r2 above this is decided in FP64
    float fk_fast;            // |F 2^32| below this: every |F 2^32 d| of a pair within rc is < 2^51
    double rc2_64;            // analytic mode: the exact test r^2 <= rc * rc
    const float *tab;
    const float *etab;        // energy pass: c0/c1 of r^-12, r^-6, r^-1, 6 floats per entry
    int off, nseg, win_lo, nwin;
    const float2 *pair_ab;
    const float2 *pair_ab_e;  // energy pass: (eps sig^12, eps sig^6) per type pair
    float u12c, u6c, u1c;     // energy pass: t12, t6 and t1 / 10 at rc^2 (the energy at rc)
    double *cell_red;         // energy pass: 11 sums per home cell
    int ntypes;
    long long *fix;
    float *f32;
    unsigned long long *counters;
    int ngroups;
    int progress_prio;        // set wave priorities from task progress (few tasks per workgroup)
    int gstart[8];
    const int *orig;          // slot -> input index (exclusions)
    const uint32_t *excl;     // exclusion windows by input index, or NULL
    BondedArgs bd;            // bonded term blocks, tasks ntasks .. ntasks + bd.nblocks - 1
};

// The exact rounding error of the staged image fl(x + S) per component (Fast2Sum with
// |S| >= |x| or S = 0): e = (x + S) - fl(x + S).  Zero where the sum is exact, which
// includes S = -L for x in [L/2, L).
__device__ __forceinline__ float4 shift_error(float4 x, float4 S)
{
    const float tx = x.x + S.x, ty = x.y + S.y, tz = x.z + S.z;
    const float zx = tx - S.x, zy = ty - S.y, zz = tz - S.z;
    return make_float4(x.x - zx, x.y - zy, x.z - zz, 0.0f);
}

// EXC: nonbonded exclusions; a queued pair whose neighbour lies in the home atom's
// exclusion window contributes no force and is not counted.
// EN (1 or 2): the energy pass instead of forces: the same pair set, each pair's energy
// from the c*_{12,6,1} tables (M/FullScale_Simulation_LJArgon.m:938-1010), its energy at rc
// and (EN = 1) its virial F d (x) d from the force tables.  Per lane and home cell the
// energies are summed in FP64 and the virial in FP32 (the lane's pairs of one cell, about
// a hundred at 0.1 atoms/A^3 and rc 12 A); from there on in FP64: lanes, waves, then cells
// in order.  One task per home cell and a fixed launch shape: bit-identical run to run.
template <bool COUL, bool F32, bool PAR, bool ANA, int APW, int TJ, bool EXC = false, int EN = 0>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_rl_halfshell(ForceArgs a)
{
    static_assert(!EN || (!F32 && !PAR && !ANA && APW == 64), "the energy pass uses the table path");
    constexpr int FSTR = COUL ? 6 : 4;   // coefficients per table entry
    constexpr int STR = FSTR;            // floats per entry of the energy pass's window
    constexpr int NACC = EN == 1 ? 8 : 2;   // EN: E, E(rc)[, W xx xy xz yy yz zz]
    constexpr int P = 64 / APW;
    static_assert(TJ <= 1024 && TJ <= NTHR, "queue entries hold 10-bit neighbour indices");
    static_assert(kBondedBlockSpan <= TI, "a bonded block accumulates in the home-tile accumulators");
    extern __shared__ float s_tab[];
    __shared__ float4 s_i[TI];
    __shared__ float4 s_j[TJ];
    __shared__ int s_it[TI];
    // neighbour tile for the candidate scan: image position relative to the home cell's
    // corner, .w = the atom's queue bits; entries [mj, mj + 64) repeat [0, 64), so the
    // rotated index of a step never wraps
    __shared__ float4 s_js[TJ + 64];
    __shared__ float4 s_je[ANA ? TJ : 1];    // analytic mode: rounding error of each staged fl(x + S)
    __shared__ unsigned long long s_iacc[EN ? 1 : 3 * TI];
    __shared__ unsigned long long s_jacc[EN ? 1 : 3 * TJ];
    __shared__ unsigned s_q[NWAVE][QCAP];
    __shared__ float2 s_ab[64];
    __shared__ int s_task;
    __shared__ unsigned long long s_stat[NWAVE][2];
    __shared__ int2 s_iom[EXC ? TI : 1];     // home atoms: input index, exclusion window
    __shared__ int s_jo[EXC ? TJ : 1];
    __shared__ int2 s_cinfo[14];             // cells of the task: first slot, atom count
    __shared__ float4 s_cshift[14];          // periodic image shift of each cell
    __shared__ int s_jslot[2][EN ? 1 : TJ];  // slot of each neighbour-tile atom (two passes)
    __shared__ double s_ered[EN ? NWAVE : 1][8];   // EN: per-wave sums
    __shared__ float s_hb[TI / 64][6];       // home tile: per-wave bounding box (corner-relative)
    __shared__ int s_wkeep[2][NWAVE];        // neighbour tile: atoms each wave keeps (two passes)

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
    const int il = lane % APW, ph = lane / APW;

    // the energy pass's window of the energy tables (the force path reads its table from
    // global memory), and the type-pair multipliers
    if (EN) {
        for (int e = tid; e < a.nwin; e += NTHR) {
            const float *p = a.etab + (size_t)(a.win_lo + e) * 6;
            float *q = s_tab + e * STR;
            q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; q[3] = p[3];
            if (FSTR == 6) { q[4] = p[4]; q[5] = p[5]; }
        }
    }
    // type-pair multipliers: A, B of the force path, or (eps sig^12, eps sig^6) of the energy
    for (int e = tid; e < a.ntypes * a.ntypes; e += NTHR) s_ab[e] = EN ? a.pair_ab_e[e] : a.pair_ab[e];
    double eacc[2] = {};   // EN: the task's E and E(rc) on this lane
    // EN 1: the virial of the task's pairs on this lane, in FP32 (six FP64 sums would push
    // the scan's registers into scratch); reduced in FP64 from the task's end on
    float vacc[EN == 1 ? 6 : 1] = {};

    float *s_iacc_f = reinterpret_cast<float *>(s_iacc);
    float *s_jacc_f = reinterpret_cast<float *>(s_jacc);
    unsigned *q = s_q[wave];
    unsigned flag = 0;
    // pairs evaluated and kept, and candidate tests of 64 lanes, by this wave (far below
    // 2^32: a launch's own totals are)
    unsigned npairs = 0, ncand = 0;
    const float rc2s = a.rc2_scan;

    // One block of bonded terms: forces accumulate in the home-tile accumulators
    // (3 x kBondedBlockSpan words: fixed point, or FP32 in FP32-atomics mode) and reach HBM
    // with one atomic per touched atom and component.
    auto bonded_task = [&](int b) {
        bonded_block(a.bd, a.bd.blocks[b], s_iacc, F32, a.fix, a.f32, tid, NTHR, flag, [] { lds_barrier(); });
    };
    // bonded tasks are compiled into the exclusion variants only (the others keep their
    // registers); the launcher runs k_bonded beside the others
    constexpr bool BND = EXC && !EN;
    const int ntot = a.ntasks + (BND ? a.bd.nblocks : 0);

    for (int it = 0;; ++it) {
        int task;
        if (EN) {
            // the energy pass: home cells dealt round-robin, one task per cell
            task = (int)blockIdx.x + it * (int)gridDim.x;
            if (task >= a.ntasks) break;
            for (int k = 0; k < 2; ++k) eacc[k] = 0.0;
            for (int k = 0; k < (EN == 1 ? 6 : 0); ++k) vacc[k] = 0.0f;
        } else if (a.tasks) {
            // the first task of each workgroup is its own index; later ones come from its
            // queue (only when there are more tasks than workgroups): queue c hands out the
            // tasks c + kQueues * k, k = (workgroups of the queue) + its counter
            int t = (int)blockIdx.x;
            if (it > 0) {
                if ((int)gridDim.x >= ntot) break;
                const int c = (int)blockIdx.x % kQueues;
                const int nb = ((int)gridDim.x - c + kQueues - 1) / kQueues;
                lds_barrier();
                if (tid == 0)
                    s_task = c + kQueues * (nb + (int)atomicAdd(reinterpret_cast<unsigned *>(&a.counters[kQueueWord + c]), 1u));
                lds_barrier();
                t = s_task;
            }
            if (t >= ntot) break;
            if (BND && t >= a.ntasks) { bonded_task(t - a.ntasks); continue; }
            task = a.tasks[t];
        } else {
            if (it > 0) break;
            if (BND && (int)blockIdx.x >= a.ntasks) { bonded_task((int)blockIdx.x - a.ntasks); continue; }
            task = blockIdx.x;
        }
        const int h = task / a.ngroups, grp = task - h * a.ngroups;
        const SysGeom &G = a.geom[a.cell_sys[h]];
        const int nc0 = G.nc[0], nc1 = G.nc[1], nc2 = G.nc[2];
        const int hl = h - G.cell_off;
        const int hz = hl % nc2, hy = (hl / nc2) % nc1, hx = hl / (nc1 * nc2);
        const int hstart = a.cell_start[h], hcount = a.cell_start[h + 1] - hstart;
        const int obeg = a.gstart[grp], oend = a.gstart[grp + 1];
        // the home cell's corner: origin of the scan's relative coordinates (the same on
        // every lane: held in scalar registers)
        auto uni = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
        const float x0 = uni((float)hx * (G.box[0] / (float)nc0)), y0 = uni((float)hy * (G.box[1] / (float)nc1)),
                    z0 = uni((float)hz * (G.box[2] / (float)nc2));
        // the scan's record of a staged neighbour: image relative to the corner, queue bits
        auto scan_rec = [&](float4 p, float4 sh, unsigned bits) {
            return make_float4((p.x + sh.x) - x0, (p.y + sh.y) - y0, (p.z + sh.z) - z0, __uint_as_float(bits));
        };

        // the task's cells: first slot, count and image shift, one thread per offset
        if (tid >= obeg && tid < oend) {
            const int o = tid;
            int rx = hx + c_hs[o][0], ry = hy + c_hs[o][1], rz = hz + c_hs[o][2];
            float4 sh = make_float4(0.f, 0.f, 0.f, 0.f);
            if (rx < 0) { rx += nc0; sh.x = -G.box[0]; } else if (rx >= nc0) { rx -= nc0; sh.x = G.box[0]; }
            if (ry < 0) { ry += nc1; sh.y = -G.box[1]; } else if (ry >= nc1) { ry -= nc1; sh.y = G.box[1]; }
            if (rz < 0) { rz += nc2; sh.z = -G.box[2]; } else if (rz >= nc2) { rz -= nc2; sh.z = G.box[2]; }
            const int c = G.cell_off + (rx * nc1 + ry) * nc2 + rz;
            const int cs = a.cell_start[c];
            s_cinfo[o] = make_int2(cs, a.cell_start[c + 1] - cs);
            s_cshift[o] = sh;
        }
        lds_barrier();

        for (int itile = 0; itile < hcount; itile += TI) {
            const int mi = min(TI, hcount - itile);
            // The first neighbour tile at or after (offset o, atom jt of its cell) that holds
            // pairs: in the home cell only pieces reaching the diagonal; cells that fit
            // are packed into one tile (not in parity mode, where the shift is per tile).
            // (the counts are read from LDS and made wave-uniform, so the tile loop stays
            // under scalar control)
            auto count_of = [&](int o) { return __builtin_amdgcn_readfirstlane(s_cinfo[o].y); };
            auto find = [&](int o, int jt) {
                for (; o < oend; ++o, jt = 0) {
                    const int cnt = count_of(o);
                    if (o == 0 && jt < itile) jt = max(jt, (itile / TJ) * TJ);   // below the diagonal
                    if (jt >= cnt) continue;
                    if (o == 0 || PAR || cnt > TJ) return NbrTile{o, o + 1, jt, min(TJ, cnt - jt)};
                    int tot = cnt, oe = o + 1;
                    while (oe < oend && tot + count_of(oe) <= TJ) tot += count_of(oe++);
                    return NbrTile{o, oe, 0, tot};
                }
                return NbrTile{oend, oend, 0, 0};
            };
            // slot of atom k of tile t and the shift of its cell
            auto locate = [&](const NbrTile &t, int k, float4 &sh) {
                int q = t.o;
                int kk = k + t.jt;
                while (q + 1 < t.oe && kk >= s_cinfo[q].y) { kk -= s_cinfo[q].y; ++q; }
                sh = s_cshift[q];
                return s_cinfo[q].x + kk;
            };
            // stage the home tile and the first neighbour tile; the neighbour tile of the
            // next pass is loaded into registers before a pass's pairs and written to LDS
            // after it, so its latency hides behind the pass
            NbrTile cur = find(obeg, 0);
            // the home tile's bounding box in the scan's corner-relative coordinates (the
            // exact minima and maxima of the scan's own home coordinates); a non-finite
            // coordinate makes it the whole space
            float blx = INFINITY, bly = INFINITY, blz = INFINITY;
            float bhx = -INFINITY, bhy = -INFINITY, bhz = -INFINITY;
            for (int k = tid; k < mi; k += NTHR) {
                const float4 p = a.pos[hstart + itile + k];
                s_i[k] = p;
                s_it[k] = a.type[hstart + itile + k];
                if (EXC) {
                    const int io = a.orig[hstart + itile + k];
                    s_iom[k] = make_int2(io, (int)a.excl[io]);
                }
                const float xr = p.x - x0, yr = p.y - y0, zr = p.z - z0;
                if (fabsf(xr) < 1e30f && fabsf(yr) < 1e30f && fabsf(zr) < 1e30f) {
                    blx = fminf(blx, xr); bly = fminf(bly, yr); blz = fminf(blz, zr);
                    bhx = fmaxf(bhx, xr); bhy = fmaxf(bhy, yr); bhz = fmaxf(bhz, zr);
                } else {
                    blx = bly = blz = -INFINITY;
                    bhx = bhy = bhz = INFINITY;
                }
            }
            if (!EN)
                for (int k = tid; k < 3 * TI; k += NTHR) s_iacc[k] = 0ull;
            // the first neighbour tile into registers
            float4 pf = make_float4(0.f, 0.f, 0.f, 0.f);
            int tf = 0, of = 0;
            if (tid < cur.mj) {
                float4 sh;
                const int sl = locate(cur, tid, sh);
                pf = a.pos[sl];
                tf = a.type[sl];
                if (EXC) of = a.orig[sl];
            }
            if (tid < TI) {
                blx = wave_min(blx); bly = wave_min(bly); blz = wave_min(blz);
                bhx = -wave_min(-bhx); bhy = -wave_min(-bhy); bhz = -wave_min(-bhz);
                if (lane == 0) {
                    s_hb[wave][0] = blx; s_hb[wave][1] = bly; s_hb[wave][2] = blz;
                    s_hb[wave][3] = bhx; s_hb[wave][4] = bhy; s_hb[wave][5] = bhz;
                }
            }
            if (!EN)
                for (int k = tid; k < 3 * TJ; k += NTHR) s_jacc[k] = 0ull;
            lds_barrier();
            blx = s_hb[0][0]; bly = s_hb[0][1]; blz = s_hb[0][2];
            bhx = s_hb[0][3]; bhy = s_hb[0][4]; bhz = s_hb[0][5];
            for (int w = 1; w < (mi + 63) / 64; ++w) {
                blx = fminf(blx, s_hb[w][0]); bly = fminf(bly, s_hb[w][1]); blz = fminf(blz, s_hb[w][2]);
                bhx = fmaxf(bhx, s_hb[w][3]); bhy = fmaxf(bhy, s_hb[w][4]); bhz = fmaxf(bhz, s_hb[w][5]);
            }
            // (the box is the same on every lane: held in scalar registers)
            blx = uni(blx); bly = uni(bly); blz = uni(blz);
            bhx = uni(bhx); bhy = uni(bhy); bhz = uni(bhz);
            // A neighbour image farther than rc from the home tile's box has no pair with
            // it: the squared distance of its scan record to the box, compared with the
            // scan's own threshold (rc^2 plus a margin far above this test's rounding and
            // the corner-relative coordinates' error against the pair's own separation), drops
            // it before the scan.  Tiles of the home cell keep every atom (the diagonal test
            // needs their own indices); a NaN coordinate gives distance 0 and is kept.
            auto near_box = [&](float4 r) {
                const float ex = fmaxf(fmaxf(blx - r.x, r.x - bhx), 0.0f);
                const float ey = fmaxf(fmaxf(bly - r.y, r.y - bhy), 0.0f);
                const float ez = fmaxf(fmaxf(blz - r.z, r.z - bhz), 0.0f);
                return !(fmaf(ez, ez, fmaf(ey, ey, ex * ex)) > rc2s);
            };
            // Stage the atoms of tile t held in registers (p, type, input index) that pass
            // the box test: the test and a ballot before a barrier (the wave's count into
            // s_wkeep[par]), the compacted records after it, at the atom's rank among the kept.
            unsigned long long kmask = 0;
            int ksl = 0;
            float4 ksh = make_float4(0.f, 0.f, 0.f, 0.f);
            auto keep_test = [&](const NbrTile &t, float4 p, int par) {
                bool keep = false;
                ksl = 0;
                ksh = make_float4(0.f, 0.f, 0.f, 0.f);
                if (tid < t.mj) {
                    ksl = locate(t, tid, ksh);
                    keep = (t.o == 0) || near_box(scan_rec(p, ksh, 0u));
                }
                kmask = __builtin_amdgcn_ballot_w64(keep);
                if (lane == 0) s_wkeep[par][wave] = (int)__popcll(kmask);
            };
            auto kept_count = [&](int par) {
                int tot = 0;
                for (int w = 0; w < NWAVE; ++w) tot += s_wkeep[par][w];
                return __builtin_amdgcn_readfirstlane(tot);
            };
            auto stage_kept = [&](float4 p, int t, int o, int par, int mk) {
                if (!((kmask >> lane) & 1ull)) return;
                int k = (int)lane_rank(kmask);
                for (int w = 0; w < wave; ++w) k += s_wkeep[par][w];
                if (ANA) s_je[k] = shift_error(p, ksh);
                const float4 r = scan_rec(p, ksh, ((unsigned)k << kJbShift) + ((unsigned)t << kTpShift));
                s_js[k] = r;
                if (k < 64) s_js[mk + k] = r;
                if (!PAR) { p.x = p.x + ksh.x; p.y = p.y + ksh.y; p.z = p.z + ksh.z; }
                s_j[k] = p;                     // .w: the scaled charge
                if (!EN) s_jslot[par][k] = ksl;
                if (EXC) s_jo[k] = o;
            };
            keep_test(cur, pf, 0);
            lds_barrier();
            int mk = kept_count(0);   // atoms of the current tile kept in LDS
            stage_kept(pf, tf, of, 0, mk);
            lds_barrier();
            int jpar = 0;

            while (cur.o < oend) {
                // Issue priority from the task's progress (3 in its first quarter of offsets,
                // 0 in its last), when workgroups hold one or two tasks each: the SIMDs issue
                // to the oldest ready wave among equal priorities, so the workgroup dispatched
                // first to a CU would otherwise run ahead and leave its partner to finish
                // alone; this keeps them in step.  (With many tasks per workgroup the queue
                // balances the load and the priorities only disturb it.)
                if (a.progress_prio) {
                    switch ((4 * (cur.o - obeg)) / (oend - obeg)) {
                    case 0: __builtin_amdgcn_s_setprio(3); break;
                    case 1: __builtin_amdgcn_s_setprio(2); break;
                    case 2: __builtin_amdgcn_s_setprio(1); break;
                    default: __builtin_amdgcn_s_setprio(0); break;
                    }
                }
                const NbrTile nxt = (cur.oe == cur.o + 1) ? find(cur.o, cur.jt + TJ) : find(cur.oe, 0);
                if (tid < nxt.mj) {
                    float4 sh;
                    const int sl = locate(nxt, tid, sh);
                    pf = a.pos[sl];
                    tf = a.type[sl];
                    if (EXC) of = a.orig[sl];
                }
                const bool self = (cur.o == 0);
                const int jtile = cur.jt, mj = mk;   // the kept atoms of the tile
                const float4 csh = s_cshift[cur.o];   // the tile's shift (single-cell tiles)
                const float sx = csh.x, sy = csh.y, sz = csh.z;

                // queue entry: home index + neighbour index << 8 + type pair << 18.  The scan
                // queues r2 <= rc2; a pair at r2 == 0 (coincident atoms, outside the
                // reference's 0 < r2 filter) or an excluded pair is dropped here, uncounted.
                auto eval = [&](unsigned e) -> bool {
                    const int ia = (int)(e & 0xffu), jb = (int)((e >> kJbShift) & kJbMask);
                    const float4 pa = s_i[ia], pb = s_j[jb];
                    float dx, dy, dz;
                    float4 je;
                    float r2;
                    if (ANA) {
                        // dx = fl((x_i - t) - e) with t = fl(x_j + S) and e its exact error:
                        // the correctly rounded separation (x_i - t is exact across a face)
                        je = s_je[jb];
                        dx = (pa.x - pb.x) - je.x;
                        dy = (pa.y - pb.y) - je.y;
                        dz = (pa.z - pb.z) - je.z;
                        r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    } else {
                        r2 = pair_geom<PAR>(pa, pb, sx, sy, sz, dx, dy, dz);
                    }
                    const float2 ab = s_ab[e >> kTpShift];
                    const float qq = pa.w * pb.w;
                    float F = 0.0f, Ee = 0.0f, Ue = 0.0f;
                    if (EN) {
                        // energy coefficients of r2's entry from the window (else the global
                        // table); the virial's force coefficients from the global table, which
                        // the caches hold (a window of both would halve the entries in LDS)
                        float c[6] = {0, 0, 0, 0, 0, 0}, ce[6] = {0, 0, 0, 0, 0, 0};
                        const unsigned wi = (__float_as_uint(r2) >> 15) - (unsigned)(a.off * 256 + a.win_lo);
                        const size_t idx = (size_t)lut_index(r2, a.off, a.nseg) * 6;
                        if (EN == 1) {
#pragma unroll
                            for (int k = 0; k < FSTR; ++k) c[k] = a.tab[idx + k];
                        }
                        if (wi < (unsigned)a.nwin) {
                            const float *t = s_tab + wi * STR;
#pragma unroll
                            for (int k = 0; k < FSTR; ++k) ce[k] = t[k];
                        } else {
#pragma unroll
                            for (int k = 0; k < FSTR; ++k) ce[k] = a.etab[idx + k];
                        }
                        if (EN == 1) F = fmaf(ab.x, fmaf(c[1], r2, c[0]), -(ab.y * fmaf(c[3], r2, c[2])));
                        Ee = fmaf(ab.x, fmaf(ce[1], r2, ce[0]), -(ab.y * fmaf(ce[3], r2, ce[2])));
                        Ue = fmaf(ab.x, a.u12c, -(ab.y * a.u6c));
                        if (COUL) {
                            const float q10 = qq * 0.1f;
                            if (EN == 1) F = fmaf(q10, fmaf(c[5], r2, c[4]), F);
                            Ee = fmaf(q10, fmaf(ce[5], r2, ce[4]), Ee);
                            Ue = fmaf(qq, a.u1c, Ue);
                        }
                    } else {
                        F = pair_scalar<COUL, PAR, ANA>(r2, ab.x, ab.y, qq, a.tab, a.off, a.nseg);
                    }
                    // the reference filter 0 < r2 <= rc2 on the pair's own r2 (the scan passed
                    // a superset): one unsigned compare of the bit patterns (r2 >= 0)
                    bool drop = ANA ? (r2 == 0.0f) : (__float_as_uint(r2) - 1u >= __float_as_uint(a.rc2));
                    if (ANA && r2 > a.rc2_lo) {
                        // within the FP32 error margin of rc^2: the exact separation
                        // D = (x_i - t) - e in FP64 (both differences exact there), and the
                        // exact test |D|^2 <= rc^2
                        const double Dx = ((double)pa.x - (double)pb.x) - (double)je.x;
                        const double Dy = ((double)pa.y - (double)pb.y) - (double)je.y;
                        const double Dz = ((double)pa.z - (double)pb.z) - (double)je.z;
                        double e2 = Dx * Dx + Dy * Dy;
                        e2 = e2 + Dz * Dz;
                        drop = drop || !(e2 <= a.rc2_64);
                    }
                    if (EXC) {
                        const int2 om = s_iom[ia];
                        const int d = s_jo[jb] - om.x + 16;
                        drop = drop || ((unsigned)d < 32u && (((unsigned)om.y >> d) & 1u));
                    }
                    const bool keep = !drop;
                    F = keep ? F : 0.0f;
                    if (EN) {
                        // the pair's energy, its energy at rc and its virial F d (x) d
                        Ee = keep ? Ee : 0.0f;
                        Ue = keep ? Ue : 0.0f;
                        eacc[0] += (double)Ee;
                        eacc[1] += (double)Ue;
                        if constexpr (EN == 1) {
                            const float gx = F * dx, gy = F * dy, gz = F * dz;
                            vacc[0] += gx * dx;
                            vacc[1] += gx * dy;
                            vacc[2] += gx * dz;
                            vacc[3] += gy * dy;
                            vacc[4] += gy * dz;
                            vacc[5] += gz * dz;
                        }
                    } else if (F32) {
                        const float fx = F * dx, fy = F * dy, fz = F * dz;
                        atomic_add_wg(&s_iacc_f[3 * ia + 0], fx);
                        atomic_add_wg(&s_iacc_f[3 * ia + 1], fy);
                        atomic_add_wg(&s_iacc_f[3 * ia + 2], fz);
                        atomic_add_wg(&s_jacc_f[3 * jb + 0], -fx);
                        atomic_add_wg(&s_jacc_f[3 * jb + 1], -fy);
                        atomic_add_wg(&s_jacc_f[3 * jb + 2], -fz);
                    } else {
                        // rint(F*d * 2^32): F is scaled first (exact while finite, and the
                        // products then carry the bits of F*d times 2^32).  With |F 2^32| below
                        // fk_fast every |F 2^32 d| of an evaluated pair (|d| <= rc) is below
                        // 2^51, where one FP64 add of 1.5 * 2^52 rounds to the integer; other
                        // lanes (rare) take the general conversion:
                        // scaled, or unscaled for a force too large to scale, with saturation
                        // and NaN flags.  Both give rint(F*d * 2^32) wherever that is in range.
                        const float Fk = F * 4294967296.0f;
                        const float xk = Fk * dx, yk = Fk * dy, zk = Fk * dz;
                        long long kx = __double_as_longlong((double)xk + 6755399441055744.0) - 0x4338000000000000LL;
                        long long ky = __double_as_longlong((double)yk + 6755399441055744.0) - 0x4338000000000000LL;
                        long long kz = __double_as_longlong((double)zk + 6755399441055744.0) - 0x4338000000000000LL;
                        if (!(fabsf(Fk) < a.fk_fast)) {
                            if (fabsf(F) < 0x1p95f) fix3_scaled(xk, yk, zk, kx, ky, kz, flag);
                            else fix3(F * dx, F * dy, F * dz, kx, ky, kz, flag);
                        }
                        atomic_add_wg(&s_iacc[3 * ia + 0], (unsigned long long)kx);
                        atomic_add_wg(&s_iacc[3 * ia + 1], (unsigned long long)ky);
                        atomic_add_wg(&s_iacc[3 * ia + 2], (unsigned long long)kz);
                        atomic_sub_wg(&s_jacc[3 * jb + 0], (unsigned long long)kx);
                        atomic_sub_wg(&s_jacc[3 * jb + 1], (unsigned long long)ky);
                        atomic_sub_wg(&s_jacc[3 * jb + 2], (unsigned long long)kz);
                    }
                    return keep;
                };

                // the wave's queue holds entries q[0, tail), tail < 64 between steps
                unsigned tail = 0;
                const int bj = mj > 0 ? il % mj : 0;
                const int nsteps = (mj + P - 1) / P;
                for (int cg0 = 0; mj > 0 && cg0 < mi; cg0 += NCHUNK * APW) {
                    if (self && itile + cg0 >= jtile + mj) break;   // rest is above the diagonal
                    const int nch = min(NCHUNK, (mi - cg0 + APW - 1) / APW);
                    // this lane's home atoms (lanes past the tile: a threshold nothing
                    // passes) and, in the home cell, the neighbour index each pair must
                    // exceed so that a pair is taken once as (lower, higher)
                    // The scan tests c_j - 2 p_i . p_j <= rc2_scan - c_i on corner-relative
                    // coordinates (c = |p|^2): each home atom holds -2 p_i and its threshold
                    // rc2_scan - c_i.  The test differs from the reference r2 <= rc2 by less
                    // than the margin folded into rc2_scan; evaluation applies the reference test.
                    float hx2[NCHUNK], hy2[NCHUNK], hz2[NCHUNK], th[NCHUNK];
                    unsigned eb[NCHUNK];
#pragma unroll
                    for (int ch = 0; ch < NCHUNK; ++ch) {
                        const int ia = cg0 + ch * APW + il;
                        if (ia < mi) {
                            const float4 p = s_i[ia];
                            const float xr = p.x - x0, yr = p.y - y0, zr = p.z - z0;
                            hx2[ch] = -2.0f * xr; hy2[ch] = -2.0f * yr; hz2[ch] = -2.0f * zr;
                            th[ch] = rc2s - fmaf(zr, zr, fmaf(yr, yr, xr * xr));
                        } else {
                            hx2[ch] = 0.0f; hy2[ch] = 0.0f; hz2[ch] = 0.0f; th[ch] = -INFINITY;
                        }
                        eb[ch] = (unsigned)ia + ((unsigned)(((ia < mi) ? s_it[ia] : 0) * a.ntypes) << kTpShift);
                    }
                    // SELF: the home cell, where a pair needs jb > itile + ia - jtile
                    // (compiled twice so the other 13 offsets skip that test)
                    const int dlim = itile - jtile;
                    auto run_steps = [&](auto self_c) {
                        constexpr bool SELF = decltype(self_c)::value;
                        for (int sft = wave; sft < nsteps; sft += NWAVE) {
                            const int k = ph + sft * P;
                            const bool kv = (P == 1) || (k < mj);
                            const float4 pj = s_js[kv ? bj + k : 0];   // bj + k < mj + 64
                            const unsigned jsh = __float_as_uint(pj.w);
                            const int jb = (int)((jsh >> kJbShift) & kJbMask);
                            const float cj = fmaf(pj.z, pj.z, fmaf(pj.y, pj.y, pj.x * pj.x));
#pragma unroll
                            for (int ch = 0; ch < NCHUNK; ++ch) {
                                if (ch < nch) {
                                    const float v = fmaf(hz2[ch], pj.z, fmaf(hy2[ch], pj.y, fmaf(hx2[ch], pj.x, cj)));
                                    const bool pass = kv && (v <= th[ch]) && (!SELF || jb > dlim + (int)(eb[ch] & 0xffu));
                                    const unsigned long long m = __builtin_amdgcn_ballot_w64(pass);
                                    // slot = tail + the lane's rank among the passing lanes
                                    unsigned *qt = q + tail;
                                    if (pass)
                                        qt[__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = eb[ch] + jsh;
                                    tail += (unsigned)__popcll(m);
                                }
                            }
                            if (tail >= 64u) {
                                // evaluate the queue's full batches of 64, then move the rest
                                // (< 64) to the front
                                unsigned head = 0;
                                do {
                                    __builtin_amdgcn_wave_barrier();
                                    npairs += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(eval(q[head + lane])));
                                    head += 64u;
                                } while (tail - head >= 64u);
                                const unsigned rest = tail - head;
                                unsigned mv = 0;
                                if ((unsigned)lane < rest) mv = q[head + lane];
                                __builtin_amdgcn_wave_barrier();
                                if ((unsigned)lane < rest) q[lane] = mv;
                                __builtin_amdgcn_wave_barrier();
                                tail = rest;
                            }
                        }
                    };
                    if (self) run_steps(std::true_type{});
                    else run_steps(std::false_type{});
                    ncand += (unsigned)((nsteps - wave + NWAVE - 1) / NWAVE) * (unsigned)nch;
                }
                if (tail) {
                    __builtin_amdgcn_wave_barrier();
                    bool kp = false;
                    if ((unsigned)lane < tail) kp = eval(q[lane]);
                    npairs += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(kp));
                    __builtin_amdgcn_wave_barrier();
                }
                keep_test(nxt, pf, jpar ^ 1);
                lds_barrier();
                const int mkn = kept_count(jpar ^ 1);
                // flush neighbour partial forces (D/Force_Write_Back_Controller.v:370-429)
                // and clear the accumulators for the next pass; the atomics stay in flight
                const int nclr = 3 * max(mj, mkn);
                if (EN) {
                    // (no forces in the energy pass)
                } else if (F32) {
                    for (int k = tid; k < nclr; k += NTHR) {
                        if (k < 3 * mj) {
                            const float v = s_jacc_f[k];
                            if (v != 0.0f) atomic_add_agent(&a.f32[4 * (size_t)s_jslot[jpar][k / 3] + k % 3], v);
                        }
                        s_jacc_f[k] = 0.f;
                    }
                } else {
                    for (int k = tid; k < nclr; k += NTHR) {
                        if (k < 3 * mj) {
                            const unsigned long long v = s_jacc[k];
                            if (v)
                                atomic_add_agent(reinterpret_cast<unsigned long long *>(a.fix + 3 * (size_t)s_jslot[jpar][k / 3] + k % 3), v);
                        }
                        s_jacc[k] = 0ull;
                    }
                }
                stage_kept(pf, tf, of, jpar ^ 1, mkn);
                jpar ^= 1;
                lds_barrier();
                cur = nxt;
                mk = mkn;
            }
            // flush reference-particle forces (D/Partial_Force_Acc.v:65-129)
            if (EN) {
                // (no forces in the energy pass)
            } else if (F32) {
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
            lds_barrier();
        }
        if constexpr (EN) {
            // the home cell's sums: lanes by a butterfly, then the waves in order; written
            // to the cell's slots (E, virial 3 x 3, E(rc)), reduced per system in cell order
#pragma unroll
            for (int k = 0; k < NACC; ++k) {
                double v = k < 2 ? eacc[k] : (double)vacc[k < 2 ? 0 : k - 2];
#pragma unroll
                for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
                if (lane == 0) s_ered[wave][k] = v;
            }
            lds_barrier();
            if (tid < 8) {
                double v = 0.0;
                if (tid < NACC)
                    for (int w = 0; w < NWAVE; ++w) v += s_ered[w][tid];
                // E, E(rc), xx xy xz yy yz zz -> slots 0, 10, 1, 2, 3, 5, 6, 9
                const int slot = tid == 0 ? 0 : tid == 1 ? 10 : tid < 5 ? tid - 1 : tid == 7 ? 9 : tid;
                double *out = a.cell_red + 11 * (size_t)h;
                out[slot] = v;
                if (tid == 3) out[4] = v;    // yx
                if (tid == 4) out[7] = v;    // zx
                if (tid == 6) out[8] = v;    // zy
            }
            lds_barrier();
        }
    }
    if (EN) return;
    // statistics (npairs/ncand are wave-uniform): one add per workgroup into its slot
    if (lane == 0) {
        s_stat[wave][0] = npairs;
        s_stat[wave][1] = (unsigned long long)ncand * 64ull;
    }
    lds_barrier();
    if (tid < 2) {
        unsigned long long v = 0;
        for (int w = 0; w < NWAVE; ++w) v += s_stat[w][tid];
        atomic_add_agent(&a.counters[8 + 2 * (blockIdx.x % kStatSlots) + tid], v);
    }
    if (flag) __hip_atomic_fetch_or(&a.counters[2], (unsigned long long)flag, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
}

// FP64 energy of every term (thread per term), for the fixed-order per-system reduction.
__global__ __launch_bounds__(256) void k_bonded_energy(BondedArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nterms) return;
    unsigned flag = 0;
    a.term_e[t] = bonded_term<true>(a, t, BondedSink{nullptr, nullptr, 0}, flag);
}

// Forces of a block of terms (one workgroup), for the launches whose force kernel does not
// take bonded tasks (no exclusion windows, FP32 atomics, parity or analytic mode): the
// terms of a block touch at most kBondedBlockSpan consecutive input indices, so they
// accumulate in LDS (fixed point, or FP32 in FP32-atomics mode) and reach HBM with one
// atomic per touched atom and component.
__global__ __launch_bounds__(256) void k_bonded(BondedArgs a)
{
    __shared__ unsigned long long s_acc[3 * kBondedBlockSpan];
    unsigned flag = 0;
    bonded_block(a, a.blocks[blockIdx.x], s_acc, a.f32 != nullptr, a.fix, a.f32, (int)threadIdx.x, (int)blockDim.x,
                 flag, [] { __syncthreads(); });
    if (flag) __hip_atomic_fetch_or(&a.counters[2], (unsigned long long)flag, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
}

static BondedArgs bonded_args(DeviceSystem &s)
{
    BondedArgs b;
    b.pos = s.pos; b.slot_of = s.slot_of; b.slot_sys = s.slot_sys; b.geom = s.geom;
    b.atoms = s.term_atoms; b.params = s.term_params; b.nterms = s.nterms;
    b.fix = (s.flags & MDRL_ACC_FP32) ? nullptr : s.fix;
    b.f32 = (s.flags & MDRL_ACC_FP32) ? s.f32 : nullptr;
    b.term_e = s.term_e; b.counters = s.counters;
    b.blocks = s.term_blocks; b.nblocks = s.nblocks;
    return b;
}

// Offset groups: the home cell is half the work of a full neighbour cell.
// Shares of a home cell's cost taken by each offset group (the self cell counts about half
// a neighbour cell).
void group_weights(int ngroups, float *w)
{
    static const float w1[1] = {13.5f}, w2[2] = {6.5f, 7.0f}, w4[4] = {3.5f, 4.0f, 3.0f, 3.0f};
    const float *v = ngroups == 1 ? w1 : ngroups == 2 ? w2 : w4;
    for (int k = 0; k < 4; ++k) w[k] = k < ngroups ? v[k] : 0.0f;
}

static void offset_groups(int ngroups, int *gstart)
{
    static const int g1[2] = {0, 14};
    static const int g2[3] = {0, 7, 14};
    static const int g4[5] = {0, 4, 8, 11, 14};
    const int *g = ngroups == 1 ? g1 : ngroups == 2 ? g2 : g4;
    for (int k = 0; k <= ngroups; ++k) gstart[k] = g[k];
}

bool schedule_supported(const DeviceSystem &s, const Schedule &sc)
{
    const bool fast = !(s.flags & (MDRL_ACC_FP32 | MDRL_EVAL_PARITY | MDRL_EVAL_ANALYTIC));
    const bool groups_ok = sc.groups == 1 || sc.groups == 2 || sc.groups == 4;
    if (!groups_ok || (sc.persistent != 0 && sc.persistent != 1)) return false;
    if (s.excl && !fast) return false;   // exclusions: the int64 table path only
    if (sc.apw == 64 && (sc.tile == 256 || sc.tile == 640)) return true;   // 640: cells packed per pass
    return fast && (sc.apw == 16 || sc.apw == 32 || sc.apw == 64) && (sc.tile == 128 || sc.tile == 256);
}

template <bool COUL, bool F32, bool PAR, bool ANA, int APW, int TJ, bool EXC = false, int EN = 0>
static void launch_rl(const ForceArgs &a0, dim3 grid, size_t lds, hipStream_t st)
{
    // Two workgroups share a CU's 160 KiB of LDS only while each stays within 80 KiB: the
    // energy pass's table window keeps its top (the bins just below rc^2, where most pairs
    // fall) and is cut from below to what fits beside this variant's static arrays; entries
    // below it are read from the global table, which gives the same bits.  (The force path
    // has no window: lds = 0.)
    static size_t static_lds = ~(size_t)0;
    if (static_lds == ~(size_t)0) {
        hipFuncAttributes fa{};
        static_lds = hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(
                         &k_rl_halfshell<COUL, F32, PAR, ANA, APW, TJ, EXC, EN>)) == hipSuccess
                         ? fa.sharedSizeBytes : 0;
    }
    ForceArgs a = a0;
    constexpr size_t kHalfCu = 80 * 1024 - 512;   // half the CU's LDS, less an allocation granule
    const size_t stride = (COUL ? 6 : 4) * sizeof(float);
    if (lds > 0 && static_lds > 0 && static_lds + lds > kHalfCu) {
        const int fit = static_lds < kHalfCu ? (int)((kHalfCu - static_lds) / stride) : 0;
        if (fit < a.nwin) {
            a.win_lo += a.nwin - fit;
            a.nwin = fit;
            lds = (size_t)fit * stride;
        }
    }
    hipLaunchKernelGGL((k_rl_halfshell<COUL, F32, PAR, ANA, APW, TJ, EXC, EN>), grid, dim3(NTHR), lds, st, a);
}

template <bool COUL, bool EXC>
static void launch_rl_fast(const ForceArgs &a, const Schedule &sc, dim3 grid, size_t lds, hipStream_t st)
{
    if (sc.tile == 640) {
        launch_rl<COUL, false, false, false, 64, 640, EXC>(a, grid, lds, st);
    } else if (sc.tile == 128) {
        if (sc.apw == 16) launch_rl<COUL, false, false, false, 16, 128, EXC>(a, grid, lds, st);
        else if (sc.apw == 32) launch_rl<COUL, false, false, false, 32, 128, EXC>(a, grid, lds, st);
        else launch_rl<COUL, false, false, false, 64, 128, EXC>(a, grid, lds, st);
    } else {
        if (sc.apw == 16) launch_rl<COUL, false, false, false, 16, 256, EXC>(a, grid, lds, st);
        else if (sc.apw == 32) launch_rl<COUL, false, false, false, 32, 256, EXC>(a, grid, lds, st);
        else launch_rl<COUL, false, false, false, 64, 256, EXC>(a, grid, lds, st);
    }
}

// the analytic, parity and FP32-accumulation variants: 64 atoms per wave, tile TJ (analytic
// takes precedence over parity, as in the pair function)
template <int TJ>
static void launch_rl_mode(const ForceArgs &a, bool coul, bool f32, bool par, bool ana, dim3 grid, size_t lds,
                           hipStream_t st)
{
    if (ana) {
        if (coul) { if (f32) launch_rl<true, true, false, true, 64, TJ>(a, grid, lds, st); else launch_rl<true, false, false, true, 64, TJ>(a, grid, lds, st); }
        else { if (f32) launch_rl<false, true, false, true, 64, TJ>(a, grid, lds, st); else launch_rl<false, false, false, true, 64, TJ>(a, grid, lds, st); }
    } else if (par) {
        if (coul) { if (f32) launch_rl<true, true, true, false, 64, TJ>(a, grid, lds, st); else launch_rl<true, false, true, false, 64, TJ>(a, grid, lds, st); }
        else { if (f32) launch_rl<false, true, true, false, 64, TJ>(a, grid, lds, st); else launch_rl<false, false, true, false, 64, TJ>(a, grid, lds, st); }
    } else {
        if (coul) launch_rl<true, true, false, false, 64, TJ>(a, grid, lds, st);
        else launch_rl<false, true, false, false, 64, TJ>(a, grid, lds, st);
    }
}

int launch_force(DeviceSystem &s, hipStream_t st)
{
    if (s.n == 0) return MDRL_OK;
    const bool f32 = s.flags & MDRL_ACC_FP32, par = s.flags & MDRL_EVAL_PARITY,
               ana = s.flags & MDRL_EVAL_ANALYTIC, coul = s.coul;
    // bonded blocks: the last tasks of the exclusion variants, else a kernel of their own
    const bool fused = s.nblocks > 0 && s.excl && !f32 && !par && !ana;
    if (s.nblocks > 0 && !fused)
        hipLaunchKernelGGL(k_bonded, dim3(s.nblocks), dim3(256), 0, st, bonded_args(s));
    ForceArgs a;
    a.bd = bonded_args(s);
    if (!fused) a.bd.nblocks = 0;
    a.pos = s.pos; a.type = s.type; a.cell_start = s.cell_start;
    a.geom = s.geom; a.cell_sys = s.cell_sys;
    a.tasks = s.sched.persistent ? s.tasks : nullptr;
    a.rc2 = s.rc2;
    a.rc2_scan = ana ? s.rc2_scan_ana : s.rc2_scan;
    // an evaluated pair has r2 <= rc2 (rc2_hi in analytic mode, below rc2_scan_ana), so each
    // |d| component is at most sqrt(rc2_scan_ana) (1 + 2^-22); 1.001 covers that and the
    // rounding of the quotient
    a.fk_fast = (float)(ldexp(1.0, 51) / (1.001 * sqrt((double)std::max(s.rc2_scan_ana, s.rc2_scan))));
    a.rc2_lo = s.rc2_lo;
    a.rc2_64 = s.rc2_64;
    a.tab = s.tab; a.off = s.exp_offset; a.nseg = s.nseg;
    a.win_lo = 0; a.nwin = 0;   // the force path reads its table from global memory
    a.pair_ab = s.pair_ab; a.ntypes = s.ntypes;
    a.fix = s.fix; a.f32 = s.f32; a.counters = s.counters;
    a.ngroups = s.sched.groups;
    offset_groups(a.ngroups, a.gstart);
    a.ntasks = s.ncells * a.ngroups;
    a.orig = s.orig; a.excl = s.excl;
    const size_t lds = 0;
    const int ntot = a.ntasks + a.bd.nblocks;
    const int grid_n = s.sched.persistent ? std::min(ntot, 2 * s.cu_count) : ntot;
    a.progress_prio = ntot <= 2 * grid_n;
    const dim3 grid(grid_n);
    const Schedule &sc = s.sched;
    if (ana || par || f32) {
        if (sc.tile == 640) launch_rl_mode<640>(a, coul, f32, par, ana, grid, lds, st);
        else launch_rl_mode<256>(a, coul, f32, par, ana, grid, lds, st);
    } else if (s.excl) {
        if (coul) launch_rl_fast<true, true>(a, sc, grid, lds, st);
        else launch_rl_fast<false, true>(a, sc, grid, lds, st);
    } else {
        if (coul) launch_rl_fast<true, false>(a, sc, grid, lds, st);
        else launch_rl_fast<false, false>(a, sc, grid, lds, st);
    }
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// binning: counting sort by cell, made stable by a per-cell rank on the input index
// (M/Simulation_Verification_LJArgon.m:143-175 keeps atoms in input order per cell).
// ---------------------------------------------------------------------------------

__device__ __forceinline__ int global_cell(const SysGeom &G, float x, float y, float z)
{
    return G.cell_off + (cell_of(x, G.inv_edge[0], G.nc[0]) * G.nc[1] + cell_of(y, G.inv_edge[1], G.nc[1])) * G.nc[2] +
           cell_of(z, G.inv_edge[2], G.nc[2]);
}

__global__ void k_assign(int n, const float4 *pos, const SysGeom *geom, const int *slot_sys,
                         int *cell, int *count, int ncells)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (s < n) {
        const float4 p = pos[s];
        c = global_cell(geom[slot_sys[s]], p.x, p.y, p.z);
        cell[s] = c;
    }
    run_add(count + count_copy() * (size_t)ncells, c, s < n);
}

// Force-task order for the persistent schedule when its tasks fit the resident grid (one
// task per workgroup, two workgroups per CU, workgroups b and b + cus sharing a CU),
// written by k_scan.  A task's cost is estimated from the populations of its home cell and
// half-shell neighbours: the candidate scan costs a pass per 64 home atoms, the evaluation
// grows with the home atoms (cost = (ceil(n/64) + 0.7 n/64) (n/2 + neighbours), times the
// offset group's share).  Tasks are counting-sorted by log2(cost) in 1/32-octave bins,
// largest first; populations change slowly, so the order is rederived at every
// kOrderEvery-th binning (counted in counters[4]).  The largest tasks go to the CUs that
// hold one workgroup, and each other CU pairs a large task with a small one.  The order
// decides only which workgroup runs a task: forces are the same for any order.
struct TaskOrder {
    const SysGeom *geom;
    const int *cell_sys;
    int *tasks;          // the persistent schedule's task list (NULL: keep the host's order)
    int ngroups;
    float w[4];          // the offset groups' shares of a cell's cost
    int pair_cus;        // CUs for the pairing placement, 0: plain order
};
constexpr int kOrderBins = 1024, kOrderEvery = 16;
// cells whose populations the order pass reads from LDS (8 per thread of k_scan); larger
// batches read them from start[]
constexpr int kOrderCells = 8192;

template <typename Count>
__device__ __forceinline__ float task_cost(Count cnt, const SysGeom &G, int c)
{
    const int k = c - G.cell_off;
    const int hz = k % G.nc[2], hy = (k / G.nc[2]) % G.nc[1], hx = k / (G.nc[1] * G.nc[2]);
    const int nh = cnt(c);
    float nb = 0.5f * (float)nh;
    for (int o = 1; o < 14; ++o) {
        int x = hx + c_hs[o][0], y = hy + c_hs[o][1], z = hz + c_hs[o][2];
        x = x < 0 ? x + G.nc[0] : (x >= G.nc[0] ? x - G.nc[0] : x);
        y = y < 0 ? y + G.nc[1] : (y >= G.nc[1] ? y - G.nc[1] : y);
        z = z < 0 ? z + G.nc[2] : (z >= G.nc[2] ? z - G.nc[2] : z);
        nb += (float)cnt(G.cell_off + (x * G.nc[1] + y) * G.nc[2] + z);
    }
    return ((float)((nh + 63) / 64) + 0.7f * (float)nh * (1.0f / 64.0f)) * nb;
}
__device__ __forceinline__ int cost_bin(float cost)
{
    // descending: bin 0 holds the largest costs
    const int b = (int)(log2f(cost + 1.0f) * 32.0f);
    return kOrderBins - 1 - min(max(b, 0), kOrderBins - 1);
}

// The force-task order of TaskOrder: one workgroup of 1024 threads; s_sum and s_wsum are
// 1024- and 16-int LDS scratch, s_cnt the populations of the first kOrderCells cells (when
// ncells <= kOrderCells; else they are read from start[]).
__device__ __forceinline__ void order_tasks(int ncells, const int *start, const TaskOrder &to, int *s_sum,
                                            int *s_wsum, const int *s_cnt)
{
    const int t = threadIdx.x;
    const bool lds = ncells <= kOrderCells;
    auto cnt = [&](int c) { return lds ? s_cnt[c] : start[c + 1] - start[c]; };
    const int g = to.ngroups;
    const int nt = ncells * g, cus = to.pair_cus;
    const int npair = nt - cus, nsingle = cus - npair;
    const bool pairing = cus > 0 && nt > cus && nt <= 2 * cus;
    s_sum[t] = 0;
    // costs of this thread's cells c = t + 1024 k, kept in registers when the cells fit LDS
    constexpr int KC = kOrderCells / 1024;
    float cst[KC];
    if (lds) {
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int c = t + 1024 * k;
            cst[k] = c < ncells ? task_cost(cnt, to.geom[to.cell_sys[c]], c) : 0.0f;
        }
    }
    __syncthreads();
    // pass 0 counts the tasks per bin, pass 1 places them
    for (int pass = 0; pass < 2; ++pass) {
        auto visit = [&](int c, float cost) {
            for (int q = 0; q < g; ++q) {
                const int bin = cost_bin(cost * to.w[q]);
                if (pass == 0) {
                    atomicAdd(&s_sum[bin], 1);
                } else {
                    const int pos = atomicAdd(&s_sum[bin], 1);
                    int at = pos;
                    if (pairing)
                        at = pos < nsingle ? npair + pos : (pos < cus ? pos - nsingle : cus + (nt - 1 - pos));
                    to.tasks[at] = c * g + q;
                }
            }
        };
        if (lds) {
#pragma unroll
            for (int k = 0; k < KC; ++k)
                if (t + 1024 * k < ncells) visit(t + 1024 * k, cst[k]);
        } else {
            for (int c = t; c < ncells; c += 1024) visit(c, task_cost(cnt, to.geom[to.cell_sys[c]], c));
        }
        __syncthreads();
        if (pass == 0) {
            // exclusive scan of the bin counts: first position of each bin
            const int lane = t & 63, w = t >> 6;
            const int v = s_sum[t];
            const int inc = wave_prefix(v);
            if (lane == 63) s_wsum[w] = inc;
            __syncthreads();
            int before = 0;
            for (int k = 0; k < w; ++k) before += s_wsum[k];
            s_sum[t] = before + inc - v;
            __syncthreads();
        }
    }
}

// The task order alone (after a schedule change), from the current cell_start.
__global__ __launch_bounds__(1024) void k_task_order(int ncells, const int *start, TaskOrder to)
{
    if (!to.tasks) return;
    __shared__ int s_sum[1024];
    __shared__ int s_wsum[16];
    __shared__ int s_cnt[kOrderCells];
    if (ncells <= kOrderCells)
        for (int c = threadIdx.x; c < ncells; c += 1024) s_cnt[c] = start[c + 1] - start[c];
    order_tasks(ncells, start, to, s_sum, s_wsum, s_cnt);
}

// Exclusive scan of the next-step counts (kCountCopies copies per cell, summed) into
// cell_start, and the start of each copy's share of each cell into the fill pointers;
// clears the counts; records the largest cell; clears the pair counters and the work
// queues of the force evaluation that follows.  One workgroup of 1024 threads walks the
// cells in tiles of 1024, one cell per thread, so every load and store is coalesced and
// each thread's kCountCopies loads are independent.  Then, for the persistent schedule,
// the force-task order (above).
__global__ __launch_bounds__(1024) void k_scan(int ncells, int *count, int *start, int *fill,
                                               unsigned long long *counters, TaskOrder to)
{
    __shared__ int s_sum[1024];
    __shared__ int s_wsum[16];
    __shared__ int s_run;
    __shared__ int s_cnt[kOrderCells];   // cell populations for the task order
    __shared__ int s_order;
    const int t = threadIdx.x;
    if (t == 0) s_run = 0;
    __syncthreads();
    int mx = 0;
    for (int base = 0; base < ncells; base += 1024) {
        const int c = base + t;
        int v[kCountCopies];
        int tot = 0;
#pragma unroll
        for (int k = 0; k < kCountCopies; ++k) {
            v[k] = (c < ncells) ? count[(size_t)k * ncells + c] : 0;
            tot += v[k];
        }
        mx = max(mx, tot);
        if (c < ncells && c < kOrderCells) s_cnt[c] = tot;
        // inclusive scan of the tile: within each wave, then across the 16 wave totals
        const int lane = t & 63, w = t >> 6;
        const int inc = wave_prefix(tot);
        if (lane == 63) s_wsum[w] = inc;
        __syncthreads();
        int before = s_run;
        for (int k = 0; k < w; ++k) before += s_wsum[k];
        if (c < ncells) {
            int r = before + inc - tot;
            start[c] = r;
#pragma unroll
            for (int k = 0; k < kCountCopies; ++k) {
                const size_t q = (size_t)k * ncells + c;
                fill[q] = r;
                r += v[k];
                count[q] = 0;
            }
        }
        __syncthreads();
        if (t == 1023) s_run = before + inc;
        __syncthreads();
    }
    // largest cell: per wave, then over the 16 waves
    mx = wave_max(mx);
    if ((t & 63) == 0) s_sum[t >> 6] = mx;
    __syncthreads();
    if (t == 0) {
        for (int k = 1; k < 16; ++k) mx = max(mx, s_sum[k]);
        start[ncells] = s_run;
        counters[3] = (unsigned long long)mx;
    }
    if (t < 2 * kStatSlots + kQueues) counters[8 + t] = 0ull;
    if (!to.tasks) return;
    // the task order, at every kOrderEvery-th binning
    if (t == 0) s_order = (int)(counters[4]++ % kOrderEvery) == 0;
    __threadfence();    // start[] written and visible to the workgroup
    __syncthreads();
    if (s_order) order_tasks(ncells, start, to, s_sum, s_wsum, s_cnt);
}

// Each slot takes the next place of its cell's share for the copy its block counted into
// (the order inside a cell is fixed afterwards by k_cell_gather).
__global__ void k_scatter(int n, const int *cell, int *fill, int *perm, int ncells)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (s < n) ? cell[s] : 0;
    const int k = run_add(fill + count_copy() * (size_t)ncells, c, s < n);
    if (s < n) perm[k] = s;
}

// Comparator t of round (k, j) of a bitonic sorting network over N = 2^p keys, all
// comparators ascending: the first round of each block size k compares an element with its
// mirror in the block, the others with the element j further.  Keys past the m real ones
// count as +infinity and never move, so comparators that reach them are skipped.
__device__ __forceinline__ void bitonic_pair(int t, int k, int j, int &lo, int &hi)
{
    const int blk = t / j, off = t - blk * j;
    if (j == (k >> 1)) { lo = blk * k + off; hi = blk * k + k - 1 - off; }
    else { lo = blk * 2 * j + off; hi = lo + j; }
}

// One workgroup per cell: orders the cell's slots along the cell's longest edge (the
// coordinate's bit pattern, ties by input index: a total order fixed by the positions alone)
// and moves their particle data into the new slots, clearing the forces and recording each
// input index's slot.  The force kernel's lanes then hold home chunks of 64 atoms that are
// slabs across that edge.  Up to 256 atoms each thread ranks its atom against the
// cell in LDS; up to SORT_CAP a bitonic network sorts (key, slot) pairs in LDS; a larger
// cell runs the same network on its slice of perm in global memory (keys read through pos
// and orig) and raises MDRL_F_CELL_OVERFLOW to say so.
constexpr int SORT_CAP = 2048, RANK_CAP = 256;
__global__ __launch_bounds__(RANK_CAP) void k_cell_gather(const int *start, int *perm, const float4 *pos,
                                                     const float4 *vel, const int *type, const int *orig,
                                                     float4 *pos2, float4 *vel2, int *type2, int *orig2,
                                                     long long *fix, float *f32, int *slot_of,
                                                     unsigned long long *counters, const SysGeom *geom,
                                                     const int *cell_sys)
{
    __shared__ int s_v[SORT_CAP];
    __shared__ long long s_k[SORT_CAP];
    const int c = blockIdx.x;
    const int b = start[c], m = start[c + 1] - b;
    const int nt = blockDim.x, tid = threadIdx.x;
    // the sort axis: the cell's longest edge (the first of equal ones)
    const SysGeom &G = geom[cell_sys[c]];
    const float ex = G.box[0] / (float)G.nc[0], ey = G.box[1] / (float)G.nc[1], ez = G.box[2] / (float)G.nc[2];
    const int ax = (ez > ex && ez > ey) ? 2 : (ey > ex ? 1 : 0);
    auto key_at = [&](float4 p, int io) {
        const float x = ax == 0 ? p.x : ax == 1 ? p.y : p.z;
        return (long long)(((unsigned long long)(unsigned)__float_as_int(x) << 32) | (unsigned)io);
    };
    auto key_of = [&](int v) { return key_at(pos[v], orig[v]); };
    auto move = [&](int s, int o) {
        const int io = orig[o];
        pos2[s] = pos[o];
        vel2[s] = vel[o];
        type2[s] = type[o];
        orig2[s] = io;
        if (slot_of) slot_of[io] = s;
        if (fix) { fix[3 * (size_t)s] = 0; fix[3 * (size_t)s + 1] = 0; fix[3 * (size_t)s + 2] = 0; }
        if (f32) reinterpret_cast<float4 *>(f32)[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    };
    if (m <= RANK_CAP) {
        // one atom per thread (RANK_CAP is the workgroup size): its data stays in registers
        // while it ranks its key against the cell's, then goes to its slot
        int io = 0, ty = 0;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f), vv = p;
        long long key = 0;
        if (tid < m) {
            const int v = perm[b + tid];
            p = pos[v];
            vv = vel[v];
            ty = type[v];
            io = orig[v];
            key = key_at(p, io);
            s_k[tid] = key;
        }
        __syncthreads();
        if (tid < m) {
            int r = 0;
            for (int u = 0; u < m; ++u) r += (s_k[u] < key);
            const int sl = b + r;
            pos2[sl] = p;
            vel2[sl] = vv;
            type2[sl] = ty;
            orig2[sl] = io;
            if (slot_of) slot_of[io] = sl;
            if (fix) { fix[3 * (size_t)sl] = 0; fix[3 * (size_t)sl + 1] = 0; fix[3 * (size_t)sl + 2] = 0; }
            if (f32) reinterpret_cast<float4 *>(f32)[sl] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        return;
    }
    int np = 1;
    while (np < m) np <<= 1;
    if (m > SORT_CAP) {
        if (tid == 0)
            __hip_atomic_fetch_or(&counters[2], 4ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int *pv = perm + b;
        for (int k = 2; k <= np; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = tid; t < (np >> 1); t += nt) {
                    int lo, hi;
                    bitonic_pair(t, k, j, lo, hi);
                    if (hi >= m) continue;
                    const int vl = pv[lo], vh = pv[hi];
                    if (key_of(vl) > key_of(vh)) { pv[lo] = vh; pv[hi] = vl; }
                }
                __threadfence();   // the round's exchanges reach every thread of the workgroup
                __syncthreads();
            }
        for (int k = tid; k < m; k += nt) move(b + k, pv[k]);
        return;
    }
    for (int k = tid; k < m; k += nt) {
        const int v = perm[b + k];
        s_v[k] = v;
        s_k[k] = key_of(v);
    }
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < (np >> 1); t += nt) {
                int lo, hi;
                bitonic_pair(t, k, j, lo, hi);
                if (hi >= m) continue;
                const long long kl = s_k[lo], kh = s_k[hi];
                if (kl > kh) {
                    s_k[lo] = kh; s_k[hi] = kl;
                    const int v = s_v[lo]; s_v[lo] = s_v[hi]; s_v[hi] = v;
                }
            }
            __syncthreads();
        }
    for (int k = tid; k < m; k += nt) move(b + k, s_v[k]);
}

int launch_bin_initial(DeviceSystem &s, hipStream_t st)
{
    if (s.n > 0)
        hipLaunchKernelGGL(k_assign, dim3(nblk(s.n, 256)), dim3(256), 0, st, s.n, s.pos, s.geom,
                           s.slot_sys, s.cell, s.cell_next, s.ncells);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

static TaskOrder task_order_args(const DeviceSystem &s)
{
    TaskOrder to{};
    to.geom = s.geom;
    to.cell_sys = s.cell_sys;
    // Only when the tasks fit the resident grid: with more tasks the work queues balance the
    // load, while the pass would add ~75 us (1 M-atom system) to ~170 us (512 replicas) to
    // every 16th binning scan; the host's list (cost per system, cells in order) is kept.
    to.tasks = s.ncells * s.sched.groups <= 2 * s.cu_count ? s.tasks : nullptr;
    to.ngroups = s.sched.groups;
    group_weights(to.ngroups, to.w);
    // the pairing placement assumes the resident grid of the launch: range-limited tasks only
    to.pair_cus = s.nblocks == 0 ? s.cu_count : 0;
    return to;
}

int launch_task_order(DeviceSystem &s, hipStream_t st)
{
    const TaskOrder to = task_order_args(s);
    if (!to.tasks || s.ncells == 0) return MDRL_OK;   // no device order for this schedule
    hipLaunchKernelGGL(k_task_order, dim3(1), dim3(1024), 0, st, s.ncells, s.cell_start, to);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

int launch_rebin(DeviceSystem &s, hipStream_t st)
{
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, s.ncells, s.cell_next, s.cell_start,
                       s.cell_fill, s.counters, task_order_args(s));
    if (s.n > 0) {
        hipLaunchKernelGGL(k_scatter, dim3(nblk(s.n, 256)), dim3(256), 0, st, s.n, s.cell,
                           s.cell_fill, s.perm, s.ncells);
        hipLaunchKernelGGL(k_cell_gather, dim3(s.ncells), dim3(RANK_CAP), 0, st, s.cell_start, s.perm,
                           s.pos, s.vel, s.type, s.orig, s.pos2, s.vel2, s.type2, s.orig2,
                           (s.flags & MDRL_ACC_FP32) ? nullptr : s.fix,
                           (s.flags & MDRL_ACC_FP32) ? s.f32 : nullptr, s.slot_of, s.counters,
                           s.geom, s.cell_sys);
    }
    MDRL_HIP(hipGetLastError());
    std::swap(s.pos, s.pos2);
    std::swap(s.vel, s.vel2);
    std::swap(s.type, s.type2);
    std::swap(s.orig, s.orig2);
    s.parity ^= 1;
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// k_update: fused Motion_Update (D/Motion_Update.v:685-805) as velocity Verlet:
//   dv  = (float)(fix * hk[type])            hk = 2^-32 * 1e-5 / m * dt/2
//   v   = vh + dv                            (full-step velocity; v = vh on the first step)
//   vh' = v + dv;  x' = wrap(x + vh'*dt);  next cell of x' counted for the rebin.
// MDRL_KICK_DRIFT gives the reference's kick-then-drift: v' = v + (dv + dv).
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
                                                const double *hk, float dt, const SysGeom *geom,
                                                const int *slot_sys, int first, int kick_drift,
                                                int *cell, int *count, int ncells,
                                                unsigned long long *counters)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned flag = 0;
    int c = 0;
    if (s < n) {
        const SysGeom &G = geom[slot_sys[s]];
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
        for (int d = 0; d < 3; ++d) {
            float dv = f3[d];
            float vn;
            if (kick_drift) {
                float dv2 = dv + dv;
                vn = vi[d] + dv2;
            } else {
                float vf = first ? vi[d] : vi[d] + dv;
                vn = vf + dv;
            }
            float step = vn * dt;
            float x = xi[d] + step;
            if (!isfinite(x) || !isfinite(vn)) flag |= 2u;
            xi[d] = wrap1(x, G.box[d]);
            vi[d] = vn;
        }
        pos[s] = make_float4(xi[0], xi[1], xi[2], p.w);
        vel[s] = make_float4(vi[0], vi[1], vi[2], 0.0f);
        c = global_cell(G, xi[0], xi[1], xi[2]);
        cell[s] = c;
    }
    run_add(count + count_copy() * (size_t)ncells, c, s < n);
    if (flag) __hip_atomic_fetch_or(&counters[2], (unsigned long long)flag, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
}

int launch_update(DeviceSystem &s, float dt, hipStream_t st)
{
    if (s.n == 0) return MDRL_OK;
    const bool f32 = s.flags & MDRL_ACC_FP32;
    hipLaunchKernelGGL(k_update, dim3(nblk(s.n, 256)), dim3(256), 0, st, s.n, s.pos, s.vel, s.type,
                       f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr, s.hk, dt, s.geom, s.slot_sys,
                       s.first_step ? 1 : 0, (s.flags & MDRL_KICK_DRIFT) ? 1 : 0, s.cell,
                       s.cell_next, s.ncells, s.counters);
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
                                                  const SysGeom *geom, const int *slot_sys,
                                                  int *cell, int *count, int ncells,
                                                  unsigned long long *counters)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned flag = 0;
    int c = 0;
    if (s < n) {
        const SysGeom &G = geom[slot_sys[s]];
        double f[3];
        for (int d = 0; d < 3; ++d)
            f[d] = f32 ? (double)f32[4 * (size_t)s + d]
                       : (double)fix[3 * (size_t)s + d] * (1.0 / 4294967296.0);
        const double fm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        const double sc = (fm * gamma > dmax) ? dmax / fm : gamma;
        float4 p = pos[s];
        float xi[3] = {p.x, p.y, p.z};
        for (int d = 0; d < 3; ++d) {
            float step = (float)(f[d] * sc);
            float x = xi[d] + step;
            if (!isfinite(x)) flag |= 2u;
            xi[d] = wrap1(x, G.box[d]);
        }
        pos[s] = make_float4(xi[0], xi[1], xi[2], p.w);
        c = global_cell(G, xi[0], xi[1], xi[2]);
        cell[s] = c;
    }
    run_add(count + count_copy() * (size_t)ncells, c, s < n);
    if (flag) __hip_atomic_fetch_or(&counters[2], (unsigned long long)flag, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
}

int launch_minimize(DeviceSystem &s, double gamma, double dmax, hipStream_t st)
{
    if (s.n == 0) return MDRL_OK;
    const bool f32 = s.flags & MDRL_ACC_FP32;
    hipLaunchKernelGGL(k_minimize, dim3(nblk(s.n, 256)), dim3(256), 0, st, s.n, s.pos,
                       f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr, gamma, dmax, s.geom,
                       s.slot_sys, s.cell, s.cell_next, s.ncells, s.counters);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// Energies.  The energy pass of k_rl_halfshell (EN): potential energy, virial and the energy
// at rc of each home cell from the c*_{12,6,1} tables (M/FullScale_Simulation_LJArgon.m:
// 938-1010), written per cell (no atomics).  k_kinetic: full-step velocity v = vh + dv(F)
// and 0.5 m v^2 per slot.
// k_sys_reduce: one workgroup per system sums its cells and slots in a fixed tree order,
// so the FP64 energies are bit-identical run to run.  Not per step: on request.
// ---------------------------------------------------------------------------------

// One workgroup per kinetic-energy chunk (at most kKeChunk consecutive slots of one system):
// full-step velocities v = vh + dv(F) per slot (written to vout when given), 0.5 m v^2
// summed by each thread over its slots in order and then by a fixed tree.
__global__ __launch_bounds__(256) void k_kinetic(const int2 *chunks, const float4 *vel, const int *type,
                                                 const long long *fix, const float *f32,
                                                 const double *hk, const double *mass,
                                                 int first, float4 *vout, double *chunk_ke)
{
    __shared__ double s_r[256];
    const int2 ch = chunks[blockIdx.x];
    double acc = 0.0;
    for (int k = threadIdx.x; k < ch.y; k += 256) {
        const int s = ch.x + k;
        const int t = type[s];
        const double hkt = hk[t];
        const float4 v = vel[s];
        float vi[3] = {v.x, v.y, v.z};
        double vv = 0.0;
        for (int d = 0; d < 3; ++d) {
            float dv = f32 ? (float)((double)f32[4 * (size_t)s + d] * 4294967296.0 * hkt)
                           : (float)((double)fix[3 * (size_t)s + d] * hkt);
            float vf = first ? vi[d] : vi[d] + dv;
            vi[d] = vf;
            vv += (double)vf * (double)vf;
        }
        const double ke = 0.5 * mass[t] * vv * 1.0e4;
        acc += ke;
        if (vout) vout[s] = make_float4(vi[0], vi[1], vi[2], (float)ke);
    }
    s_r[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s_r[threadIdx.x] += s_r[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_ke[blockIdx.x] = s_r[0];
}

// One workgroup per system: its kinetic-energy chunks, its cells' sums (energy, virial,
// energy at rc) and its bonded terms' energies, each thread over a fixed stride and then a
// fixed tree, so the FP64 energies are bit-identical run to run.
constexpr int kRedThreads = 1024;
__global__ __launch_bounds__(kRedThreads) void k_sys_reduce(const SysGeom *geom, const double *cell_red,
                                                            const double *chunk_ke, const double *term_e,
                                                            const int *term_off, double *red)
{
    __shared__ double s_r[kRedStride][kRedThreads / 64];
    const SysGeom &G = geom[blockIdx.x];
    const int ncell = G.nc[0] * G.nc[1] * G.nc[2];
    double acc[kRedStride] = {};
    for (int k = threadIdx.x; k < G.ke_n; k += kRedThreads) acc[0] += chunk_ke[G.ke_off + k];
    for (int c = threadIdx.x; c < ncell; c += kRedThreads) {
        for (int k = 0; k < 10; ++k) acc[1 + k] += cell_red[11 * (size_t)(G.cell_off + c) + k];
        acc[12] += cell_red[11 * (size_t)(G.cell_off + c) + 10];
    }
    if (term_e)
        for (int t = term_off[blockIdx.x] + threadIdx.x; t < term_off[blockIdx.x + 1]; t += kRedThreads)
            acc[11] += term_e[t];
    // within each wave by a butterfly, then over the waves in order
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < kRedStride; ++k) {
        double v = acc[k];
        for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
        if (lane == 0) s_r[k][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < kRedStride) {
        double v = 0.0;
        for (int q = 0; q < kRedThreads / 64; ++q) v += s_r[threadIdx.x][q];
        red[kRedStride * (size_t)blockIdx.x + threadIdx.x] = v;
    }
}

int launch_energy(DeviceSystem &s, float4 *vout, hipStream_t st, int mode)
{
    if (s.n > 0) {
        const bool f32 = s.flags & MDRL_ACC_FP32;
        hipLaunchKernelGGL(k_kinetic, dim3(s.nke), dim3(256), 0, st, s.ke_chunks, s.vel, s.type,
                           f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr, s.hk, s.mass,
                           s.first_step ? 1 : 0, vout, s.chunk_ke);
        if (mode == kEnergyKinetic) {   // full-step velocities only
            MDRL_HIP(hipGetLastError());
            return MDRL_OK;
        }
        // the energy pass: every home cell with all 14 offsets, 640-atom neighbour tiles, cells
        // dealt round-robin to a fixed grid (so the sums do not depend on the force schedule)
        ForceArgs a{};
        a.pos = s.pos; a.type = s.type; a.cell_start = s.cell_start;
        a.geom = s.geom; a.cell_sys = s.cell_sys;
        a.tasks = nullptr;
        a.ntasks = s.ncells;
        a.rc2 = s.rc2; a.rc2_scan = s.rc2_scan;
        a.tab = s.tab; a.etab = s.etab; a.off = s.exp_offset; a.nseg = s.nseg;
        a.win_lo = s.win.lo; a.nwin = s.win.n;
        a.pair_ab = s.pair_ab; a.pair_ab_e = s.pair_ab_e; a.ntypes = s.ntypes;
        a.u12c = s.u12c; a.u6c = s.u6c; a.u1c = s.u1c;
        a.cell_red = s.cell_red;
        a.counters = s.counters;
        a.ngroups = 1;
        offset_groups(1, a.gstart);
        a.orig = s.orig; a.excl = s.excl;
        const int grid_n = std::min(s.ncells, 2 * s.cu_count);
        a.progress_prio = s.ncells <= 2 * grid_n;
        const bool vir = mode == kEnergyVirial;
        const size_t lds = (size_t)s.win.n * (s.coul ? 6 : 4) * sizeof(float);
        const dim3 g(grid_n);
        if (vir) {
            if (s.coul) {
                if (s.excl) launch_rl<true, false, false, false, 64, 640, true, 1>(a, g, lds, st);
                else launch_rl<true, false, false, false, 64, 640, false, 1>(a, g, lds, st);
            } else {
                if (s.excl) launch_rl<false, false, false, false, 64, 640, true, 1>(a, g, lds, st);
                else launch_rl<false, false, false, false, 64, 640, false, 1>(a, g, lds, st);
            }
        } else {
            if (s.coul) {
                if (s.excl) launch_rl<true, false, false, false, 64, 640, true, 2>(a, g, lds, st);
                else launch_rl<true, false, false, false, 64, 640, false, 2>(a, g, lds, st);
            } else {
                if (s.excl) launch_rl<false, false, false, false, 64, 640, true, 2>(a, g, lds, st);
                else launch_rl<false, false, false, false, 64, 640, false, 2>(a, g, lds, st);
            }
        }
        if (s.nterms > 0)
            hipLaunchKernelGGL(k_bonded_energy, dim3(nblk(s.nterms, 256)), dim3(256), 0, st, bonded_args(s));
    }
    hipLaunchKernelGGL(k_sys_reduce, dim3(s.nsys), dim3(kRedThreads), 0, st, s.geom, s.cell_red, s.chunk_ke,
                       s.nterms > 0 ? s.term_e : nullptr, s.term_off, s.red);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// fills (kernels instead of hipMemsetAsync, so a captured step stays valid)
// ---------------------------------------------------------------------------------

__global__ void k_zero_forces(int n, long long *fix, float *f32, unsigned long long *counters)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < 2 * kStatSlots + kQueues) counters[8 + s] = 0ull;
    if (s >= n) return;
    if (fix) { fix[3 * (size_t)s] = 0; fix[3 * (size_t)s + 1] = 0; fix[3 * (size_t)s + 2] = 0; }
    if (f32) reinterpret_cast<float4 *>(f32)[s] = make_float4(0.f, 0.f, 0.f, 0.f);
}

int launch_zero_forces(DeviceSystem &s, hipStream_t st)
{
    const bool f32 = s.flags & MDRL_ACC_FP32;
    hipLaunchKernelGGL(k_zero_forces, dim3(nblk(std::max(s.n, 2 * kStatSlots + kQueues), 256)), dim3(256), 0, st, s.n,
                       f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr, s.counters);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

// ---------------------------------------------------------------------------------
// summation hook: forces from a provider outside the range-limited path
// ---------------------------------------------------------------------------------
// The reference adds short-range, long-range and bonded forces per particle before
// Motion_Update (Summation_Logic/Summation_Logic.v:3-6, Force_Summation_Unit.v:62-126);
// its long-range unit exchanges {x, y, z, q} records for force records
// (Long_Range/Top_Electro.sv:3-20).  Here the provider sees input order, so it needs no
// knowledge of the cell sort: k_lr_export writes the records and clears the provider's
// outputs, k_lr_import adds the provider's forces into the slot accumulators.

// rint(f * 10 * 2^32) as int64, saturated at +-2^62: kJ/mol/A to the fixed-point unit.
// The product is exact in FP64 (24-bit mantissa times 5 * 2^33), so one rounding.
__device__ __forceinline__ long long fix_of_kj_per_a(float f, unsigned &flag)
{
    const double x = (double)f * 42949672960.0;
    if (fabs(x) < 0x1p51) return __double_as_longlong(x + 6755399441055744.0) - 0x4338000000000000LL;
    if (x != x) { flag |= 2u; return 0; }
    if (x > 0x1p62) { flag |= 1u; return 1LL << 62; }
    if (x < -0x1p62) { flag |= 1u; return -(1LL << 62); }
    return (long long)rint(x);
}

__global__ __launch_bounds__(256) void k_lr_export(int n, int nsys, const float4 *__restrict__ pos,
                                                   const int *__restrict__ orig, const float *__restrict__ q,
                                                   float4 *__restrict__ xq, float *__restrict__ f,
                                                   double *__restrict__ e)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nsys) e[k] = 0.0;
    if (k >= n) return;
    const int i = orig[k];
    const float4 p = pos[k];
    xq[i] = make_float4(p.x, p.y, p.z, q[i]);
    f[3 * (size_t)i] = 0.0f;
    f[3 * (size_t)i + 1] = 0.0f;
    f[3 * (size_t)i + 2] = 0.0f;
}

__global__ __launch_bounds__(256) void k_lr_import(int n, const int *__restrict__ orig, const float *__restrict__ f,
                                                   long long *__restrict__ fix, float *__restrict__ f32,
                                                   unsigned long long *counters)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = orig[k];
    const float fx = f[3 * (size_t)i], fy = f[3 * (size_t)i + 1], fz = f[3 * (size_t)i + 2];
    if (f32) {
        float4 a = reinterpret_cast<float4 *>(f32)[k];
        a.x += fx * 10.0f;
        a.y += fy * 10.0f;
        a.z += fz * 10.0f;
        reinterpret_cast<float4 *>(f32)[k] = a;
        return;
    }
    unsigned flag = 0;
    fix[3 * (size_t)k] += fix_of_kj_per_a(fx, flag);
    fix[3 * (size_t)k + 1] += fix_of_kj_per_a(fy, flag);
    fix[3 * (size_t)k + 2] += fix_of_kj_per_a(fz, flag);
    if (flag) __hip_atomic_fetch_or(&counters[2], (unsigned long long)flag, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
}

int launch_lr_export(DeviceSystem &s, const float *q, float4 *xq, float *f, double *e, hipStream_t st)
{
    const int m = std::max(s.n, s.nsys);
    if (m == 0) return MDRL_OK;
    hipLaunchKernelGGL(k_lr_export, dim3(nblk(m, 256)), dim3(256), 0, st, s.n, s.nsys, s.pos, s.orig, q, xq, f, e);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

int launch_lr_import(DeviceSystem &s, const float *f, hipStream_t st)
{
    if (s.n == 0) return MDRL_OK;
    const bool f32 = s.flags & MDRL_ACC_FP32;
    hipLaunchKernelGGL(k_lr_import, dim3(nblk(s.n, 256)), dim3(256), 0, st, s.n, s.orig, f,
                       f32 ? nullptr : s.fix, f32 ? s.f32 : nullptr, s.counters);
    MDRL_HIP(hipGetLastError());
    return MDRL_OK;
}

}  // namespace mdrl
