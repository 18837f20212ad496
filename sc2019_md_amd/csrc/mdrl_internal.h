// mdrl_internal.h -- host-side structures shared by the C ABI (mdrl_api.cpp) and the
// kernel launchers (mdrl_kernels.hip).  Not installed; not part of the ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/mdrl.h"

struct md_rl_table {
    int exp_offset = 121;
    int nseg = 14;
    int nent = 0;                       // nseg * 256
    bool has[MDRL_NTERMS] = {};
    std::vector<float> c0[MDRL_NTERMS], c1[MDRL_NTERMS];
};

namespace mdrl {

void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);   // records the error, returns MDRL_EHIP

#define MDRL_HIP(call)                                             \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return ::mdrl::hip_fail(_e, #call);  \
    } while (0)

// Fixed-point scale: forces accumulate as int64 multiples of 2^-32 kJ/mol/nm.
constexpr double kFixScale = 4294967296.0;

// Geometry of one periodic system (one replica).
struct Geometry {
    int n = 0;
    int nc[3] = {1, 1, 1};
    int ncells = 1;
    float box[3] = {0, 0, 0};
    float inv_edge[3] = {0, 0, 0};      // (float)(nc / box), the binning rule
    float rc2 = 0;
};

// Table window kept in LDS by the force kernel: entries [lo, lo + n).
struct TableWindow {
    int lo = 0, n = 0;
    int stride = 6;                     // floats per entry (6 with r^-3, 4 without)
};

// Device state of one system.
struct DeviceSystem {
    Geometry g;
    int ntypes = 1;
    int flags = 0;
    bool coul = false;
    int cell_cap = 0;
    // particle data in cell-sorted order (double-buffered for the rebin gather)
    float4 *pos = nullptr, *pos2 = nullptr;   // x, y, z, scaled charge
    float4 *vel = nullptr, *vel2 = nullptr;   // vx, vy, vz, kinetic energy
    int *type = nullptr, *type2 = nullptr;
    int *orig = nullptr, *orig2 = nullptr;    // input index of each slot
    int *cell = nullptr;                      // cell of each slot after the update
    int *perm = nullptr;                      // new slot -> old slot
    int *cell_start = nullptr;                // ncells + 1
    int *cell_next = nullptr;                 // ncells: counts of the next binning
    int *cell_fill = nullptr;                 // ncells
    long long *fix = nullptr;                 // n x 3, int64 fixed point
    float *f32 = nullptr;                     // n x 4, FP32 atomics mode
    float *tab = nullptr;                     // force table, 6 floats / entry
    float *etab = nullptr;                    // energy table (12, 6, 1), 6 floats / entry
    float2 *pair_ab = nullptr;                // ntypes^2 (A, B) for the kernel's mode
    float2 *pair_ab_e = nullptr;              // ntypes^2 (A, B) energy multipliers
    double *hk = nullptr;                     // ntypes: half-kick factor
    double *mass = nullptr;                   // ntypes
    unsigned long long *counters = nullptr;   // [0] pairs [1] candidates [2] flags [3] max cell
    double *red = nullptr;                    // [0] KE [1] PE [2..10] virial
    int exp_offset = 121, nseg = 14;
    TableWindow win;
    bool first_step = true;
};

// Kernel launchers (mdrl_kernels.hip).
int launch_bin_initial(DeviceSystem &s, hipStream_t st);   // cell ids of the current slots
int launch_rebin(DeviceSystem &s, hipStream_t st);         // scan, scatter, sort, gather
int launch_force(DeviceSystem &s, hipStream_t st);
int launch_update(DeviceSystem &s, float dt, hipStream_t st);
int launch_minimize(DeviceSystem &s, double gamma, double dmax, hipStream_t st);
int launch_energy(DeviceSystem &s, hipStream_t st);
int launch_zero_forces(DeviceSystem &s, hipStream_t st, bool forces);  // pair counters (+ forces)
int launch_zero_red(DeviceSystem &s, hipStream_t st);     // FP64 reductions
int launch_kinetic(DeviceSystem &s, float4 *vout, hipStream_t st);
int launch_eval_pairs(const float *d_tab, int exp_offset, int nseg, const float *ref,
                      const float *nbr, const float *pa, const float *pb, const float *pqq,
                      float *fs, float *fxyz, int64_t n, float rc2, int flags, hipStream_t st);

struct CModelDev {
    int n;
    float *x, *v, *f, *e, *ke;   // x,v,f: 3 x n SoA; e, ke: n
    float *sums;                 // 2 floats per iteration
};
int launch_cmodel_iteration(CModelDev &m, const md_rl_cmodel_cfg &cfg, int it,
                            hipStream_t st);

}  // namespace mdrl
