// mdrl_kernels.hip -- gfx950 kernels of the pair-batch API and of the C model.
//
// Built with -ffp-contract=off: every rounding happens where it is written.
//
// Reference paths: D/ = Range_Limited_demo/SourceCode/, R/ = Range_Limited/,
// M/ = D/MatlabScripts/, O/ = D/OpenCL_HDL_Lib/.
#include <hip/hip_runtime.h>

#include "mdrl_device.h"

namespace mdrl {

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

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
        const float *c = tab + (size_t)lut_index(r2, off, nseg) * 6;
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

// Plain streaming copy (16 B per lane per iteration), for calibrating byte counters.
__global__ void k_stream_copy(float4 *__restrict__ dst, const float4 *__restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int launch_stream_copy(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (n == 0) return MDRL_OK;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, st, static_cast<float4 *>(dst),
                       static_cast<const float4 *>(src), n);
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
