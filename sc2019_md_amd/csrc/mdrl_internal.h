// mdrl_internal.h -- host-side structures shared by the C ABI (mdrl_api.cpp) and the
// kernel launchers (mdrl_system.hip, mdrl_kernels.hip).  Not installed; not part of the ABI.
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
    uint64_t id = 0;                    // unique per table (tables are immutable)
};

namespace mdrl {

void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);   // records the error, returns MDRL_EHIP

#define MDRL_HIP(call)                                             \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return ::mdrl::hip_fail(_e, #call);  \
    } while (0)

// A memory image read from a Quartus .mif or an Intel HEX file (mdrl_cells.cpp): depth
// words of `width` bits (32 or 96), each as width / 32 parts, most significant first;
// given[a] says whether the file assigned word a.
struct WordFile {
    int width = 0;
    int depth = 0;
    std::vector<uint32_t> parts;
    std::vector<uint8_t> given;
};
int read_words(const char *path, WordFile &w);   // MDRL_OK, or an error recorded

// Fixed-point scale: forces accumulate as int64 multiples of 2^-32 kJ/mol/nm.
constexpr double kFixScale = 4294967296.0;

// One periodic system of a batch (a single system is a batch of one).  Its atoms occupy
// slots [atom_off, atom_off + n) and its cells [cell_off, cell_off + nc0*nc1*nc2) of the
// batch-wide arrays; sorting never moves an atom out of its system's range.
struct SysGeom {
    float box[3];
    float inv_edge[3];      // (float)(nc / box): the binning rule
    int nc[3];
    int cell_off;
    int atom_off;
    int n;
    int ke_off, ke_n;       // the system's kinetic-energy chunks (kKeChunk slots each)
};

// With MDRL_COARSE_CELLS the default grid holds at least this many atoms per cell on
// average (or one cell).
constexpr int kMinCellAtoms = 128;

// Kinetic energies are reduced in chunks of at most kKeChunk slots within one system (one
// workgroup each, a fixed tree), then per system over its chunks.
constexpr int kKeChunk = 2048;

// Force-kernel schedule (autotuned): home atoms per wavefront, neighbour-tile depth,
// half-shell offset groups per workgroup, persistent work queue.
struct Schedule {
    int apw = 64;
    int tile = 640;
    int groups = 1;
    int persistent = 1;
};

// Table window kept in LDS by the force kernel: entries [lo, lo + n).
struct TableWindow {
    int lo = 0, n = 0;
    int stride = 6;                     // floats per entry (6 with r^-3, 4 without)
};

// Pair and candidate counts are spread over kStatSlots counter pairs (workgroup b adds
// into slot b mod kStatSlots), so thousands of workgroups do not serialise on one address.
constexpr int kStatSlots = 64;
// The persistent force kernel drains its task list through kQueues counters: queue c
// serves the tasks c, c + kQueues, ... to the workgroups b with b mod kQueues = c.
constexpr int kQueues = 8;
constexpr int kQueueWord = 8 + 2 * kStatSlots;
constexpr int kCounterWords = kQueueWord + kQueues;

// Copies of the per-cell counts of the binning (see count_copy in the kernels).
constexpr int kCountCopies = 8;

constexpr int kRedStride = 13;          // per-system FP64 reductions: KE, PE, virial[9], bonded PE, PE at rc

// Device state of a batch of systems.
struct DeviceSystem {
    int nsys = 1;
    int n = 0;                                // atoms over all systems
    int ncells = 0;                           // cells over all systems
    float rc2 = 0.f;
    // analytic mode's exact cutoff: the scan passes r2 <= rc2_hi (FP32); pairs with
    // r2 > rc2_lo are then decided by the exact FP64 r^2 <= rc2_64 = rc * rc
    float rc2_hi = 0.f, rc2_lo = 0.f;
    double rc2_64 = 0.0;
    // the candidate scan's threshold on its |p_i|^2 + |p_j|^2 - 2 p_i.p_j (corner-relative
    // coordinates): the cutoff of the mode plus the bound on that r2's error
    float rc2_scan = 0.f, rc2_scan_ana = 0.f;
    // the energy tables' t12, t6 and t1 / 10 at rc^2: the pair energy at the cutoff
    float u12c = 0.f, u6c = 0.f, u1c = 0.f;
    int ntypes = 1;
    int flags = 0;
    bool coul = false;
    std::vector<SysGeom> h_geom;
    SysGeom *geom = nullptr;                  // [nsys]
    int *cell_sys = nullptr;                  // [ncells] system of each cell
    int *slot_sys = nullptr;                  // [n] system of each slot (static)
    int *tasks = nullptr;                     // force tasks, largest first
    int ntasks = 0;
    // particle data in cell-sorted order (double-buffered for the rebin gather)
    float4 *pos = nullptr, *pos2 = nullptr;   // x, y, z, scaled charge
    float4 *vel = nullptr, *vel2 = nullptr;   // vx, vy, vz (half step), -
    int *type = nullptr, *type2 = nullptr;
    int *orig = nullptr, *orig2 = nullptr;    // input index of each slot
    int *cell = nullptr;                      // cell of each slot after the update
    int *perm = nullptr;                      // new slot -> old slot
    int *cell_start = nullptr;                // ncells + 1
    int *cell_next = nullptr;                 // kCountCopies x ncells: counts of the next binning
    int *cell_fill = nullptr;                 // kCountCopies x ncells: fill pointers
    long long *fix = nullptr;                 // n x 3, int64 fixed point
    float *f32 = nullptr;                     // n x 4, FP32 atomics mode
    float *tab = nullptr;                     // force table, 6 floats / entry
    float *etab = nullptr;                    // energy table (12, 6, 1), 6 floats / entry
    float2 *pair_ab = nullptr;                // ntypes^2 (A, B) for the kernel's mode
    float2 *pair_ab_e = nullptr;              // ntypes^2 (eps sig^12, eps sig^6)
    double *hk = nullptr;                     // ntypes: half-kick factor
    double *mass = nullptr;                   // ntypes
    unsigned long long *counters = nullptr;   // kCounterWords: [2] flags [3] max cell [4] binnings, then
                                              // kStatSlots x {pairs, candidates}, then the
                                              // kQueues work counters
    double *cell_red = nullptr;               // ncells x 11: PE, virial, PE at rc per home cell
    int2 *ke_chunks = nullptr;                // nke: first slot, slot count of each chunk
    double *chunk_ke = nullptr;               // nke: kinetic energy per chunk
    int nke = 0;
    double *red = nullptr;                    // nsys x kRedStride
    // bonded terms and nonbonded exclusions (batch-wide input indices)
    uint32_t *excl = nullptr;                 // n exclusion windows by input index, or NULL
    int *slot_of = nullptr;                   // n: input index -> slot, kept by the rebin
    int nterms = 0;
    int4 *term_atoms = nullptr;               // nterms: bond (i,j,-1,-1), angle (i,j,k,-1), torsion
    float4 *term_params = nullptr;            // nterms: kernel constants (factor 10 folded in)
    double *term_e = nullptr;                 // nterms: FP64 energy per term (on request)
    int *term_off = nullptr;                  // nsys + 1: terms of each system
    int4 *term_blocks = nullptr;              // per block: first term, end, first atom, span
    int nblocks = 0;
    int exp_offset = 121, nseg = 14;
    TableWindow win;
    Schedule sched;
    int cu_count = 256;
    bool first_step = true;
    int parity = 0;                           // ping-pong state of the double buffers
};

// Kernel launchers (mdrl_system.hip).
int launch_bin_initial(DeviceSystem &s, hipStream_t st);   // cell ids of the current slots
int launch_rebin(DeviceSystem &s, hipStream_t st);         // scan, scatter, sort, gather
int launch_task_order(DeviceSystem &s, hipStream_t st);    // force-task order from cell_start
int launch_force(DeviceSystem &s, hipStream_t st);
int launch_update(DeviceSystem &s, float dt, hipStream_t st);
int launch_minimize(DeviceSystem &s, double gamma, double dmax, hipStream_t st);
int launch_zero_forces(DeviceSystem &s, hipStream_t st);   // forces and pair counters
// Energy reductions red[] per system: kinetic only (full-step velocities into vout), the
// potential energy and energy at rc, or those with the pair virial.
enum { kEnergyKinetic = 0, kEnergyPotential = 1, kEnergyVirial = 2 };
int launch_energy(DeviceSystem &s, float4 *vout, hipStream_t st, int mode);
bool schedule_supported(const DeviceSystem &s, const Schedule &sc);
void group_weights(int ngroups, float *w);       // offset groups' shares of a cell's cost
// Summation hook: {x, y, z, q} records in input order out of the slots (clearing the
// provider's forces f, n x 3, and energies e, nsys), and the provider's forces back in.
int launch_lr_export(DeviceSystem &s, const float *q, float4 *xq, float *f, double *e, hipStream_t st);
int launch_lr_import(DeviceSystem &s, const float *f, hipStream_t st);
constexpr int kBondedBlockSpan = 256;     // input indices one block of bonded terms may touch
constexpr int kBondedBlockTerms = 1024;   // terms per block

// Kernel launchers (mdrl_kernels.hip).
int launch_eval_pairs(const float *d_tab, int exp_offset, int nseg, const float *ref,
                      const float *nbr, const float *pa, const float *pb, const float *pqq,
                      float *fs, float *fxyz, int64_t n, float rc2, int flags, hipStream_t st);

int launch_stream_copy(void *dst, const void *src, size_t bytes, hipStream_t st);

struct CModelDev {
    int n;
    float *x, *v, *f, *e, *ke;   // x,v,f: 3 x n SoA; e, ke: n
    float *sums;                 // 2 floats per iteration
};
int launch_cmodel_iteration(CModelDev &m, const md_rl_cmodel_cfg &cfg, int it,
                            hipStream_t st);

}  // namespace mdrl
