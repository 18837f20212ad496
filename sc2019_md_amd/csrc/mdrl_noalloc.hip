// mdrl_noalloc.hip -- md_rl_eval_pairs_ex_noalloc: the pair-batch evaluation of
// md_rl_eval_pairs_ex on caller-owned device buffers, for callers that cannot allocate
// device memory (code under graph capture, serving loops, harnesses that forbid it).
//
// The one buffer the pair kernel needs beyond its inputs and outputs is the interleaved
// table (6 floats per entry: c0/c1 of r^-14, r^-8, r^-3), which md_rl_eval_pairs_ex stages
// into a fresh allocation on every call.  Here it is staged into `scratch`, a file-scope
// device array of fixed size that each GPU holds until the process ends: the table's
// columns are copied into the first part, and k_interleave writes the 6-float rows the
// kernel reads into the second.  The table staged last is remembered per device and is not
// staged again while calls keep using it.  Every call on a GPU shares that GPU's array, so
// calls that overlap on one GPU (several host threads, or other streams still running a
// call with a different table) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mdrl_internal.h"

namespace {

constexpr int kMaxEntries = 32 * 256;                 // tables of up to 32 segments
constexpr size_t kColFloats = 6 * (size_t)kMaxEntries;  // staged columns
constexpr size_t kScratchFloats = 2 * kColFloats;       // columns + interleaved rows: 384 KiB

static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
float *g_scratch[kMaxDevices];       // per device: the address of that device's scratch
uint64_t g_staged[kMaxDevices];      // per device: id of the table its scratch holds (0: none)
std::mutex g_mu;

// rows[e * 6 + k] = cols[k * nent + e] for the terms in mask, zero for the others
__global__ void k_interleave(const float *cols, float *rows, int nent, unsigned mask)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    for (int k = 0; k < 6; ++k) rows[6 * (size_t)e + k] = ((mask >> (k / 2)) & 1u) ? cols[(size_t)k * nent + e] : 0.0f;
}

// The device's scratch address, looked up once per device with that device current.
int scratch_of(int dev, float **out)
{
    if (dev < 0 || dev >= kMaxDevices) {
        mdrl::set_error("md_rl_eval_pairs_ex_noalloc: device %d beyond the %d this library tracks", dev, kMaxDevices);
        return MDRL_EINVAL;
    }
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return mdrl::hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = static_cast<float *>(p);
    }
    *out = g_scratch[dev];
    return MDRL_OK;
}

}  // namespace

extern "C" int md_rl_eval_pairs_ex_noalloc(const md_rl_table *t, const float *ref_xyz,
                                           const float *nbr_xyz, const float *p_a,
                                           const float *p_b, const float *p_qq, float *force_s,
                                           float *force_xyz, int64_t npairs, float rc2,
                                           int flags, void *hip_stream)
{
    if (!t || npairs < 0 || (npairs > 0 && (!ref_xyz || !nbr_xyz || !force_xyz))) {
        mdrl::set_error("md_rl_eval_pairs_ex_noalloc: bad argument");
        return MDRL_EINVAL;
    }
    if (t->nent <= 0 || t->nent > kMaxEntries) {
        mdrl::set_error("md_rl_eval_pairs_ex_noalloc: table of %d entries exceeds the %d entries "
                        "(%zu KiB) of the staging scratch", t->nent, kMaxEntries,
                        sizeof(float) * kScratchFloats / 1024);
        return MDRL_EINVAL;
    }
    if ((flags & MDRL_PAIR_COULOMB) && !t->has[MDRL_TERM_3]) {
        mdrl::set_error("md_rl_eval_pairs_ex_noalloc: Coulomb requested but the table has no r^-3 term");
        return MDRL_EINVAL;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        mdrl::set_error("no HIP device visible");
        return MDRL_ENODEV;
    }
    if (npairs == 0) return MDRL_OK;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    int dev = 0;
    MDRL_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mu);
    float *ptr = nullptr;
    int rc = scratch_of(dev, &ptr);
    if (rc) return rc;
    float *cols = ptr, *rows = ptr + kColFloats;
    if (g_staged[dev] != t->id) {
        // the table's columns straight from its host arrays, then the rows; waits for the
        // copies, so the caller may free the table as soon as the call returns
        g_staged[dev] = 0;
        const int terms[3] = {MDRL_TERM_14, MDRL_TERM_8, MDRL_TERM_3};
        unsigned mask = 0;
        for (int k = 0; k < 3; ++k) {
            if (!t->has[terms[k]]) continue;
            mask |= 1u << k;
            MDRL_HIP(hipMemcpyAsync(cols + (size_t)(2 * k) * t->nent, t->c0[terms[k]].data(),
                                    sizeof(float) * t->nent, hipMemcpyHostToDevice, st));
            MDRL_HIP(hipMemcpyAsync(cols + (size_t)(2 * k + 1) * t->nent, t->c1[terms[k]].data(),
                                    sizeof(float) * t->nent, hipMemcpyHostToDevice, st));
        }
        hipLaunchKernelGGL(k_interleave, dim3((t->nent + 255) / 256), dim3(256), 0, st, cols, rows,
                           t->nent, mask);
        MDRL_HIP(hipGetLastError());
        MDRL_HIP(hipStreamSynchronize(st));
        g_staged[dev] = t->id;
    }
    return mdrl::launch_eval_pairs(rows, t->exp_offset, t->nseg, ref_xyz, nbr_xyz, p_a, p_b, p_qq,
                                   force_s, force_xyz, npairs, rc2, flags, st);
}
