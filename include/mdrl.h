/*
 * mdrl.h -- C ABI of libmdrl, the MI355X (gfx950) range-limited force path.
 *
 * Plain C: no C++ types, no exceptions and no device types cross this boundary.
 * Every entry point returns an int status (MDRL_OK = 0, negative MDRL_E* on error;
 * md_rl_last_error() gives the message of the calling thread's last failure), except
 * MD_Evaluation_Model(), which keeps the reference's "return 1" convention.
 *
 * Ownership: the caller owns every host array it passes; the library copies what it
 * needs and owns all device memory inside opaque handles.  A handle is not
 * thread-safe; distinct handles may be used from distinct threads.  Calls are
 * synchronous unless the name ends in _async; *_async calls take a hipStream_t
 * passed as void* (NULL = the handle's own stream).
 *
 * Units on the system path: positions in Angstrom, velocities in Angstrom/fs, forces
 * in kJ/mol/Angstrom, masses in amu, energies in kJ/mol, time step in fs, charges in e.
 *
 * Reference paths: D/ = Range_Limited_demo/SourceCode/, R/ = Range_Limited/,
 * O/ = Range_Limited_demo/SourceCode/OpenCL_HDL_Lib/, M/ = D/MatlabScripts/.
 */
#ifndef MDRL_H
#define MDRL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDRL_VERSION 1

/* ---- status ---------------------------------------------------------------- */
enum {
    MDRL_OK = 0,
    MDRL_EINVAL = -1,      /* bad argument */
    MDRL_ENOMEM = -2,      /* host or device allocation failed */
    MDRL_EIO = -3,         /* file missing or malformed */
    MDRL_EHIP = -4,        /* HIP runtime error */
    MDRL_ECELL = -5,       /* cell geometry: cell edge < rc, or a cell over capacity */
    MDRL_ENONFINITE = -6,  /* NaN/Inf position, velocity or force */
    MDRL_ESATURATED = -7,  /* a fixed-point force contribution was clamped */
    MDRL_ENODEV = -8,      /* no gfx950 device visible */
};

/* Message for the calling thread's most recent failing call ("" if none). */
const char *md_rl_last_error(void);
int md_rl_version(void);

/* ---- interpolation tables ---------------------------------------------------- */
/* First-order tables over r^2: entry index = {exp8(r2) - exp_offset, mantissa[22:15]}
 * (256 bins per octave), term(r2) = c1*r2 + c0.  Replaces the lut*_* ROMs initialised
 * from c{0,1}_{14,8,3}.mif/.hex (D/lut0_14.v:34-38, D/RL_LJ_Evaluate_Pairs_1st_Order.v:
 * 176-199) and the MATLAB writers (M/LJ_Coulomb_no_smooth_poly_interpolation_function.m).
 * ApoA1 layout: exp_offset 121, 14 segments; LJArgon layout: 125, 9 segments. */
enum {
    MDRL_TERM_14 = 0, MDRL_TERM_8 = 1, MDRL_TERM_3 = 2,   /* force terms */
    MDRL_TERM_12 = 3, MDRL_TERM_6 = 4, MDRL_TERM_1 = 5,   /* energy terms */
    MDRL_NTERMS = 6
};
typedef struct md_rl_table md_rl_table;

/* Reads c0_<k>.<ext> / c1_<k>.<ext> for k in 14,8,3,12,6,1 from dir; ext is tried in
 * the order mif, hex, txt.  The 14 and 8 terms are required, the others optional. */
int md_rl_table_load(md_rl_table **out, const char *dir, int exp_offset, int nseg);
/* Builds a table from caller arrays (nseg*256 entries each; NULL = term absent). */
int md_rl_table_from_arrays(md_rl_table **out, int exp_offset, int nseg,
                            const float *const c0[MDRL_NTERMS],
                            const float *const c1[MDRL_NTERMS]);
/* Reads one table file (Quartus .mif, Intel .hex with 4-byte records, or MATLAB .txt
 * with one decimal per line) into out[cap]; returns the entry count or < 0. */
int md_rl_table_read_file(const char *path, float *out, int cap);
int md_rl_table_get(const md_rl_table *t, int term, const float **c0, const float **c1,
                    int *nentries, int *exp_offset, int *nseg);
void md_rl_table_free(md_rl_table *t);
/* Directory of the ApoA1-layout tables shipped with the library. */
const char *md_rl_default_table_dir(void);
/* Loads the shipped ApoA1 tables (offset 121, 14 segments). */
int md_rl_table_load_default(md_rl_table **out);

/* ---- cell-memory initialisation files ------------------------------------------------ */
/* The FPGA initialises one position memory per cell from cell_ini_file_X_Y_Z.{mif,hex}
 * (X, Y, Z from 1; D/CellMemoryModules/{ApoA1,LJArgon}_Mem_Ini_File/,
 * LJArgon_v_File_64_Cells/), written by M/GenInputCellList_LJArgon.m:140-205 and
 * M/GenInputCellList_ApoA1.m:138-203: 96-bit words, word 0 = {0, 0, count}, words
 * 1..count = {z, y, x} big-endian IEEE singles, zero words up to DEPTH.  The format
 * follows the path's extension: Quartus .mif (as the generator prints it, count in %X)
 * or Intel HEX (12-byte records, as Quartus converts the .mif).  Flags: MDRL_FILE_CRLF
 * ends lines with CR LF, as the shipped files do. */
enum { MDRL_FILE_CRLF = 1 };
/* Reads one cell file into xyz[cap][3] (x, y, z); returns the particle count or < 0. */
int md_rl_cell_file_read(const char *path, float *xyz, int cap);
/* Writes count particles (xyz[count][3]) as a cell file of depth words (depth > count). */
int md_rl_cell_file_write(const char *path, const float *xyz, int count, int depth, int flags);
/* The Motion_Update boundary memory cell_ini_boundary_value (M/GenCellBoundary_LJArgon.m:
 * 48-60): 32-bit words, word k = k * edge.  Read returns the word count. */
int md_rl_boundary_file_read(const char *path, float *bounds, int cap);
int md_rl_boundary_file_write(const char *path, float edge, int depth, int flags);
/* Reads the cell files <dir>/<prefix>X_Y_Z.hex (else .mif) for lo[d] <= X, Y, Z <= hi[d]
 * (prefix NULL = "cell_ini_file_") in the generator's order, X outermost, into particle
 * rows [cap][11] (positions in cols 0-2, the rest zero; md_rl_set_particles' layout) and,
 * if cell is not NULL, each particle's (X, Y, Z) into cell[cap][3]; returns the total. */
int md_rl_cell_dir_read(const char *dir, const char *prefix, const int lo[3], const int hi[3],
                        float *particle, int32_t *cell, int cap);

/* ---- pair batch ---------------------------------------------------------------- */
/* Evaluates independent (reference, neighbour) pairs.  Replaces the OpenCL kernel
 * LJ(ref_id, neighbor_id, ref_x..z, neighbor_x..z, Force_out_x..z)
 * (O/test_hdl_lib/device/LJ.cl:31-80) and its RTL function
 * float4 RL_LJ_Evaluation(int2 particle_id, float4 reference_pos, float4 neighbor_pos)
 * (O/RL_LJ_Evaluation.h:1, O/HDL_Source/RL_LJ_Evaluation_OpenCL_Top.v:94-134):
 * d = ref - nbr, r2 = |d|^2, F = t14 - t8 (unit-coefficient tables), out = F*d, and
 * 0 for pairs with r2 == 0 or r2 > rc2 (O/HDL_Source/RL_LJ_Evaluate_Pairs_1st_Order.v:
 * 150-152).  ref_id/nbr_id are carried as in the reference and may be NULL.
 * Host pointers; synchronous. */
enum {
    MDRL_PAIR_PARITY = 1,   /* unfused r2 and c1*r2+c0, as the MATLAB golden model */
    MDRL_PAIR_STRICT = 2,   /* accept r2 < rc2 (MATLAB) instead of r2 <= rc2 (HDL) */
    MDRL_PAIR_COULOMB = 4,  /* add p_qq * t3 (RL_Evaluate_Pairs_1st_Order.v) */
};
int md_rl_eval_pairs(const md_rl_table *t, const int32_t *ref_id, const int32_t *nbr_id,
                     const float *ref_x, const float *ref_y, const float *ref_z,
                     const float *nbr_x, const float *nbr_y, const float *nbr_z,
                     float *force_x, float *force_y, float *force_z, int64_t npairs,
                     float rc2, int flags);
/* As md_rl_eval_pairs with the per-pair multipliers of D/RL_Evaluate_Pairs_1st_Order.v:
 * 53-55 (p_a, p_b, p_qq; NULL = 1, 1, 0): F = p_a*t14 - p_b*t8 + p_qq*t3.  force_s
 * (may be NULL) receives the scalar F. */
int md_rl_eval_pairs_ex(const md_rl_table *t, const float *ref_xyz, const float *nbr_xyz,
                        const float *p_a, const float *p_b, const float *p_qq,
                        float *force_s, float *force_xyz, int64_t npairs, float rc2,
                        int flags);
/* md_rl_eval_pairs_ex on device buffers, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  All pointers except t are
 * device pointers (p_a, p_b, p_qq, force_s may be NULL as above); the work is enqueued on
 * hip_stream (NULL: the null stream) and the call returns without waiting for it.  The
 * interleaved table lives in a fixed scratch array of 384 KiB per GPU (tables of at most
 * 32 segments; larger ones are rejected with MDRL_EINVAL), staged from t's host arrays
 * when t differs from the table staged last on this GPU; that staging waits for its
 * copies (so a call that stages cannot be captured into a graph: make one call with the
 * table before capturing).  Calls that overlap on one GPU share its scratch and race. */
int md_rl_eval_pairs_ex_noalloc(const md_rl_table *t, const float *ref_xyz,
                                const float *nbr_xyz, const float *p_a, const float *p_b,
                                const float *p_qq, float *force_s, float *force_xyz,
                                int64_t npairs, float rc2, int flags, void *hip_stream);

/* ---- system path ------------------------------------------------------------------ */
/* Half-shell cell list (D/Particle_Pair_Gen_HalfShell_64_Cells.v), r2 filter
 * (D/Filter_Logic.v:234), table force (D/RL_LJ_Evaluate_Pairs_1st_Order.v), Newton's-
 * third-law accumulation (D/Partial_Force_Acc.v, D/Force_Write_Back_Controller.v) and
 * Motion_Update (D/Motion_Update.v), on the C model's particle layout
 * float particle[N][11] (R/C_Model_ForceEval/MD_Evaluation_Model.cpp:35):
 * 0-2 position, 3-5 velocity, 6-8 force, 9 potential energy, 10 kinetic energy. */
enum {
    MDRL_ACC_FP32 = 1,       /* FP32 atomics instead of int64 fixed point (not deterministic) */
    MDRL_EVAL_ANALYTIC = 2,  /* FP32 direct r^-14/r^-8/r^-3 instead of the tables */
    MDRL_EVAL_PARITY = 4,    /* unfused arithmetic, d = (xi - xj) - image (reference order) */
    MDRL_KICK_DRIFT = 8,     /* reference integrator: v += a*dt; x += v*dt (no half kicks) */
    MDRL_NO_COULOMB = 16,    /* drop the r^-3 term even if charges are set */
    MDRL_COARSE_CELLS = 32,  /* default grid coarsened while cells average < 128 atoms: faster
                                for sparse systems; the pair set is the same, but a pair across
                                the periodic box may be taken from its other atom, so its force
                                can differ in the last bits from the floor(box/rc) grid's */
};

typedef struct md_rl_params {
    int n;              /* atoms */
    float box[3];       /* periodic box edges (A) */
    float rc;           /* cutoff (A) */
    int ncell[3];       /* cells per axis; 0 = floor(box/rc) (with MDRL_COARSE_CELLS, then one
                           fewer on the axis of those with the most cells, ties to the shorter
                           edge, while the average cell holds fewer than 128 atoms). Single
                           systems only; batches use 0. Cell edge must be >= rc */
    float dt;           /* time step (fs) */
    int ntypes;         /* atom types (1..8) */
    const float *eps;   /* ntypes: LJ epsilon (kJ/mol), Lorentz-Berthelot mixing */
    const float *sigma; /* ntypes: LJ sigma (A) */
    const float *mass;  /* ntypes: mass (amu) */
    float coulomb_k;    /* Coulomb constant (kJ A / mol / e^2); 0 = no r^-3 term */
    int flags;          /* MDRL_ACC_* | MDRL_EVAL_* | MDRL_KICK_DRIFT | ... */
    const md_rl_table *table;  /* NULL = the shipped ApoA1 tables */
    int cell_capacity;  /* max atoms per cell; 0 = automatic */
    int reserved[7];
} md_rl_params;

typedef struct md_rl_ctx md_rl_ctx;

typedef struct md_rl_stats {
    int64_t pair_evals;       /* in-cutoff unordered pairs of the last force evaluation */
    int64_t candidates;       /* half-shell candidate pairs tested */
    int max_cell_count;       /* largest cell at the last rebin */
    int flags;                /* sticky MDRL_F_* bits raised by the kernels */
    double force_ms;          /* mean device time of the force kernel over the timed steps
                                 of the last synchronised batch (HIP events on its stream) */
    double step_ms;           /* mean device time of a full step over the same steps */
    int timed_steps;          /* steps behind the two means */
} md_rl_stats;
enum { MDRL_F_SATURATED = 1, MDRL_F_NONFINITE = 2, MDRL_F_CELL_OVERFLOW = 4 };

int md_rl_create(md_rl_ctx **out, const md_rl_params *p);
/* Positions (cols 0-2) and velocities (cols 3-5) of n particles in the N x 11 layout;
 * positions are wrapped into the box. */
int md_rl_set_particles(md_rl_ctx *c, const float *particle, int n);
/* Type index (0..ntypes-1) and charge (e) per particle; charge may be NULL. */
int md_rl_set_types(md_rl_ctx *c, const int32_t *type, const float *charge, int n);
/* Bonded terms of system sys (0 for a single system), replacing any set before; they
 * replace the streams of Bonded/Bonded_Top.v:202-380.  atoms[nterms][4] are the system's
 * input indices with -1 padding: a bond (i, j, -1, -1), an angle (i, j, k, -1) with
 * vertex j, or a torsion (i, j, k, l).  params[nterms][4], positions in A:
 *   bond    (r0, k [kJ/mol/A^2]):            U = k (r - r0)^2           (Bonded/calc_bond_v2.v:83-110)
 *   angle   (theta0 [rad], k [kJ/mol/rad^2]): U = k (theta - theta0)^2  (Bonded/calc_angle_v2.v:113-147)
 *   torsion (k [kJ/mol], n in 1..6, phase [rad], -): U = k (1 + cos(n phi + phase))
 *                                                                     (Bonded/calcTorsion_v2.v:70-274)
 * The atoms of one term must lie within 256 consecutive input indices.  FP32 forces
 * join the range-limited forces in the same fixed-point buffer (and every
 * md_rl_force_eval / step / minimize); their FP64 energy joins the potential energy. */
int md_rl_set_bonded(md_rl_ctx *c, int sys, int nterms, const int32_t *atoms, const float *params);
/* Nonbonded exclusions of system sys (no counterpart in the reference, whose golden
 * models exclude by distance, M/FullScale_Simulation_LJArgon.m:801): bit d of mask[i]
 * removes the pair (i, i + d - 16) (system input indices) from the range-limited forces,
 * energies and pair counts.  Windows must be symmetric with bit 16 clear; NULL removes
 * them.  Needs the default int64 table path. */
int md_rl_set_exclusions(md_rl_ctx *c, int sys, const uint32_t *mask);
/* Host-only check of exclusion windows as md_rl_set_exclusions applies it (no counterpart
 * in the reference): every set bit d of mask[i] names an atom i + d - 16 in [0, n) whose
 * own window has bit 32 - d set; bit 16 (the atom itself) and bit 0 (atom i - 16, whose
 * reciprocal bit 32 lies outside the window) are rejected.  Returns MDRL_OK or
 * MDRL_EINVAL with the offending atom in md_rl_last_error(). */
int md_rl_check_exclusions(const uint32_t *mask, int n);
/* Summation hook for forces computed outside this library (a long-range solver, an
 * applied field, restraints).  The reference sums short-range, long-range and bonded
 * forces per particle before Motion_Update (Summation_Logic/Summation_Logic.v:3-6,
 * Force_Summation_Unit.v:62-126), its long-range unit taking {x, y, z, q} records and
 * returning force records (Long_Range/Top_Electro.sv:3-20).  After the range-limited and
 * bonded forces of every force evaluation, step and minimisation, the library writes the
 * records of the current positions in input order, clears force and energy, calls
 * fn(user, view, stream) and adds view->force into the accumulators: each component is
 * converted like a pair force, rint(f * 10 * 2^32) in int64 (FP32 mode: f * 10 added in
 * FP32).  md_rl_energy / md_rl_system_energy call fn with want_energy = 1, add
 * view->energy to the potential energy and discard the forces.
 * fn must enqueue its work on the stream it is given (a hipStream_t as void*) and not
 * synchronise the device: with md_rl_set_graph on, the calls made while a step is
 * captured are replayed by every later launch of that graph, so the work must depend on
 * the view's buffers only (their addresses stay fixed).  A nonzero return aborts the
 * calling entry point with MDRL_EINVAL.  fn = NULL removes the hook. */
typedef struct md_rl_lr_view {
    int nsys;                  /* systems in the context */
    int n;                     /* atoms over all systems */
    const int32_t *atom_off;   /* host, nsys + 1: system k holds input indices [atom_off[k], atom_off[k+1]) */
    const float *box;          /* host, nsys x 3 (A) */
    const float *xq;           /* device, n x 4: x, y, z (A, in [0, box)), charge (e); input order */
    float *force;              /* device, n x 3: the provider's forces (kJ/mol/A), input order; zero on entry */
    double *energy;            /* device, nsys: the provider's energy per system (kJ/mol); zero on entry */
    int want_energy;           /* 1 when called for md_rl_energy / md_rl_system_energy */
} md_rl_lr_view;
typedef int (*md_rl_lr_fn)(void *user, const md_rl_lr_view *view, void *hip_stream);
int md_rl_set_long_range(md_rl_ctx *c, md_rl_lr_fn fn, void *user);
/* Forces at the current positions (cols 6-8 after md_rl_get_particles). */
int md_rl_force_eval(md_rl_ctx *c);
/* nsteps velocity-Verlet (or kick-drift) steps: force, Motion_Update, rebin. */
int md_rl_step(md_rl_ctx *c, int nsteps);
int md_rl_step_async(md_rl_ctx *c, int nsteps, void *hip_stream);
int md_rl_sync(md_rl_ctx *c);
/* Copies back positions, velocities (full step), forces and per-atom kinetic energy. */
int md_rl_get_particles(md_rl_ctx *c, float *particle, int n);
/* Capped steepest descent on the current positions (velocities untouched): each atom
 * moves by F * min(gamma, max_disp / |F|) per iteration (F in kJ/mol/nm, gamma in
 * A per kJ/mol/nm, max_disp in A), then rebin and forces.  Pre-relaxes jittered lattices
 * whose overlaps would otherwise blow up the first steps. */
int md_rl_minimize(md_rl_ctx *c, int iterations, float max_disp, float gamma);
/* Raw int64 fixed-point forces (2^-32 kJ/mol/nm units), n x 3, input order. */
int md_rl_get_forces_fixed(md_rl_ctx *c, int64_t *fix, int n);
/* FP64 energy of the current positions from the c*_{12,6,1} tables (potential),
 * kinetic energy of the full-step velocities and the pair virial (sum r_ij F_ij), each
 * reduced in a fixed order so repeated calls give identical bits. */
int md_rl_energy(md_rl_ctx *c, double *pe, double *ke, double virial[9]);
/* The parts of md_rl_energy for system sys (-1: all systems), in kJ/mol.  cutoff_shift is
 * the sum over the in-cutoff, non-excluded pairs of the table pair energy at rc: the
 * reference's potentials are truncated without a shift (the *_no_smooth tables), so the
 * potential jumps by a pair's energy at rc whenever the pair crosses the cutoff, and
 * potential - cutoff_shift (the shifted potential, whose forces are the same) is what
 * velocity Verlet conserves together with the kinetic energy.  No reference counterpart. */
typedef struct md_rl_energy_terms {
    double kinetic;            /* 0.5 m v^2 at full-step velocities */
    double potential;          /* range_limited + bonded + long_range, as md_rl_energy */
    double range_limited;      /* pair energies from the c*_{12,6,1} tables */
    double bonded;
    double long_range;         /* the summation hook's energy */
    double virial[9];          /* sum over pairs of F_ij (x) r_ij, kJ/mol */
    double cutoff_shift;
} md_rl_energy_terms;
int md_rl_get_energy_terms(md_rl_ctx *c, int sys, md_rl_energy_terms *out);
int md_rl_get_stats(md_rl_ctx *c, md_rl_stats *s);
/* Enables HIP-event timing of each force kernel and full step (off by default). */
int md_rl_set_timing(md_rl_ctx *c, int on);
int md_rl_destroy(md_rl_ctx *c);

/* ---- batches of independent systems -------------------------------------------------- */
/* Many periodic systems of uneven sizes stepped together: every kernel of a step is one
 * launch over all systems (the force kernel optionally as a persistent grid draining a
 * largest-first task queue), and the step can be captured in a hipGraph.  Systems share
 * rc, dt, flags, species and table (common); each has its own atom count and box.  The
 * md_rl_* calls above apply to the whole batch (md_rl_energy sums over systems); the
 * single-system calls md_rl_set_particles/set_types/get_particles/get_forces_fixed
 * require a batch of one. */
int md_rl_create_batch(md_rl_ctx **out, const md_rl_params *common, int nsys,
                       const int32_t *n, const float *box /* nsys x 3 */);
int md_rl_num_systems(const md_rl_ctx *c);
int md_rl_system_size(const md_rl_ctx *c, int sys, int32_t *n, float box[3], int ncell[3]);
/* Stages one system's particles (n x 11: positions, velocities) and/or types and charges
 * (either may be NULL); the next step, force evaluation or read-back uploads and bins
 * every staged system.  md_rl_commit uploads at once. */
int md_rl_set_system(md_rl_ctx *c, int sys, const float *particle, const int32_t *type,
                     const float *charge);
int md_rl_commit(md_rl_ctx *c);
int md_rl_get_system(md_rl_ctx *c, int sys, float *particle);
int md_rl_get_system_forces_fixed(md_rl_ctx *c, int sys, int64_t *fix);
int md_rl_system_energy(md_rl_ctx *c, int sys, double *pe, double *ke, double virial[9]);

/* Capture md_rl_step's steps in hipGraphs (16 steps per graph launch, then pairs of steps); off
 * by default. */
int md_rl_set_graph(md_rl_ctx *c, int on);
/* Captures, without running them, the graphs that md_rl_step / md_rl_step_async of nsteps
 * steps would replay from the current state, so that those calls only replay graphs (no
 * capture or instantiation inside them).  Needs md_rl_set_graph(c, 1) and a context that
 * has taken its first step since its particles were set (that step runs eagerly);
 * otherwise it captures nothing and returns MDRL_OK.  No reference counterpart. */
int md_rl_graph_prepare(md_rl_ctx *c, int nsteps);

/* Force-kernel schedule.  atoms_per_wave: home atoms per 64-lane wavefront (16, 32 or
 * 64; below 64, 64/apw lanes share a home atom and split its neighbours); tile_depth:
 * neighbour atoms staged in LDS per pass (128 or 256, or 640 with 64 atoms per wave;
 * neighbour cells that fit are packed into one pass); offset_groups: workgroups per home
 * cell, splitting its 14 half-shell offsets (1, 2 or 4); persistent: a fixed grid draining
 * a largest-first task queue instead of one workgroup per task.  Default {64, 640, 1, 1}.
 * MDRL_ACC_FP32, MDRL_EVAL_PARITY and MDRL_EVAL_ANALYTIC take tiles 256 and 640 at 64
 * atoms per wave only (MDRL_EINVAL otherwise).
 * All schedules give bit-identical int64 forces. */
typedef struct md_rl_schedule {
    int atoms_per_wave;
    int tile_depth;
    int offset_groups;
    int persistent;
} md_rl_schedule;
int md_rl_set_schedule(md_rl_ctx *c, const md_rl_schedule *s);
int md_rl_get_schedule(const md_rl_ctx *c, md_rl_schedule *s);
/* Times every available schedule on the current positions (reps force evaluations
 * each, HIP events) and keeps the fastest. */
int md_rl_autotune(md_rl_ctx *c, int reps, md_rl_schedule *best, double *best_ms);
/* md_rl_autotune through a cache of winning schedules keyed by device, library build and
 * workload shape (atoms, systems, cells, types, rc, flags, Coulomb, exclusions, bonded
 * terms): a hit applies the cached schedule (and evaluates forces, as a sweep leaves them)
 * without sweeping; a miss sweeps and records.  The cache lives in the process and in the
 * JSON Lines file $MDRL_TUNE_CACHE ("" or "off": none; default
 * $XDG_CACHE_HOME/mdrl/schedules.jsonl or ~/.cache/mdrl/schedules.jsonl).  hit (may be
 * NULL) receives 1 on a cache hit, 0 after a sweep.  No reference counterpart (the
 * reference's FPGA schedule is fixed at synthesis). */
int md_rl_autotune_cached(md_rl_ctx *c, int reps, md_rl_schedule *best, double *best_ms, int *hit);

/* ---- the C model entry point ------------------------------------------------------- */
/* Drop-in for int MD_Evaluation_Model() (R/C_Model_ForceEval/MD_Evaluation_Model.h:1):
 * reads the PDB named by $MDRL_INPUT (default: the reference's ar_gas.pdb path), runs
 * the reference's 100 kick-then-drift iterations of all-pairs LJ on the GPU with the
 * reference's arithmetic, prints one line per iteration in the reference's format
 * (:159) and returns 1; 0 on error. */
int MD_Evaluation_Model(void);

typedef struct md_rl_cmodel_cfg {
    int n;                 /* particles (reference: 500) */
    int iterations;        /* reference: 100 */
    float eps, sigma, mass, dt, rc;   /* reference literals, :14-24 */
    float box[3];          /* :26-28 */
    int print;             /* print the per-iteration line */
} md_rl_cmodel_cfg;
/* Fills the reference's constants with its promotions (:11-29). */
void md_rl_cmodel_default(md_rl_cmodel_cfg *cfg);
/* Runs the C model on the GPU over particle[n][11] in place (positions and velocities
 * in, everything out).  energies (may be NULL) receives 2 floats per iteration: the
 * printed system energy and kinetic energy. */
int md_rl_cmodel_run(const md_rl_cmodel_cfg *cfg, float *particle, float *energies);

/* ---- synthetic systems ----------------------------------------------------------------- */
/* Seeded generators of the benchmark systems (SplitMix64-seeded xoshiro256**), shared by
 * the CPU baselines and the GPU so both see identical inputs. */
typedef struct md_rl_system {
    int n, ntypes;
    float box[3];
    float rc;
    float dt;
    float coulomb_k;
    float *particle;   /* n x 11 (pos, vel; other columns 0) */
    int32_t *type;     /* n */
    float *charge;     /* n */
    float eps[8], sigma[8], mass[8];
    /* bonded topology (workload 5; 0 / NULL otherwise), in md_rl_set_bonded's format */
    int nterms;
    int32_t *term_atoms;   /* nterms x 4 */
    float *term_params;    /* nterms x 4 */
    uint32_t *excl;        /* n exclusion windows, or NULL */
} md_rl_system;
/* which: 1 = 512-atom argon; 2 = 92,224-atom LJ+Coulomb; 3 = 1,048,576-atom;
 * 4 = one replica of the batched set (index = replica, 0..511);
 * 5 = one replica of the bonded set (index = replica, 0..127): 64 linear chains of 48
 *     united-atom beads at 0.033 beads/A^3 (45.33 A box, rc 10 A, dt 2 fs), bonds
 *     r0 1.53 A, angles 111 deg, torsions n = 3, charges +-0.25 e alternating, LJ and
 *     Coulomb excluded within 3 bonds. */
int md_rl_make_system(md_rl_system *s, int which, int index, uint64_t seed);
void md_rl_free_system(md_rl_system *s);

/* ---- diagnostics --------------------------------------------------------------------- */
/* No counterpart in the reference.  Copies bytes (a multiple of 16) from device buffer
 * src to device buffer dst with one plain streaming kernel and waits for it: a kernel of
 * exactly known traffic (bytes read, bytes written) for scaling hardware byte counters. */
int md_rl_stream_copy(void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
