/*
 * oracle.h -- CPU restatement of the reference's range-limited force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libmdrl, its kernels or its
 * host code) includes, links or calls this code.  Only tests/, the smoke check and
 * bench.py's CPU-baseline leg load it, and there only as the reference results or as the
 * timed CPU baseline.
 *
 * Every function cites the reference file:line it restates.  Path prefixes:
 *   D/ = Range_Limited_demo/SourceCode/
 *   R/ = Range_Limited/
 *   M/ = Range_Limited_demo/SourceCode/MatlabScripts/
 *   O/ = Range_Limited_demo/SourceCode/OpenCL_HDL_Lib/
 *
 * Pinning (what checks each function against something other than itself):
 *   or_lut_index, or_pair_golden     -> M/VERIFICATION_REFERENCE_OUTPUT.txt, bit-exact
 *   or_bin_ceil, or_pair_stream_text -> M/VERIFICATION_PARTICLE_PAIR_INPUT.txt, SHA-256
 *   or_cmodel_run                    -> the reference C model itself, compiled from
 *                                       R/C_Model_ForceEval under oracle/_ref, printed
 *                                       energies compared line by line
 *   or_rl_forces_fix                 -> or_fp64_direct (closed form per pair, brute
 *                                       force over images), N3L sum = 0 exactly,
 *                                       two-atom closed forms; bit-exact vs the GPU
 *   or_fp64_direct                   -> two-atom closed forms (LJ minimum at
 *                                       2^(1/6) sigma, Coulomb k q q / r^2),
 *                                       brute-force vs cell-list agreement
 *                                       (energy and virial); virial trace ==
 *                                       -dU/dl under scaling by l
 *   or_motion_update,                -> closed forms: free flight, constant force
 *   or_minimize_update
 *   or_fix                           -> exact values: 1 -> 2^32, ties to even,
 *                                       +-1e30 -> +-2^62
 *   or_table_fp64                    -> within 1e-5 of or_rl_forces_fix, within 1e-4
 *                                       of or_fp64_direct (the tables' own error)
 *   or_cmodel_force_sample           -> two-atom closed form of the C model's force
 *   or_cell_of                       -> through pair counts == brute force over all
 *                                       images (any consistent rule with cells >= rc)
 *   or_atan2f                        -> FP64 arctan2 on 4,000 points (< 5e-7), axes
 *   or_bonded_fix, or_bonded_fp64    -> central finite differences of the FP64 energy,
 *                                       single-term closed forms, minimum image
 *   exclusion windows                -> brute-force count of non-excluded pairs and an
 *                                       independent FP64 sum that skips them
 */
#ifndef MDRL_ORACLE_H
#define MDRL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tables --------------------------------------------------------------- */
/* Index of the first-order table entry for r2: {exp8(r2) - offset, mantissa[22:15]},
 * segment forced to 0 when out of [0, nseg).  D/RL_LJ_Evaluate_Pairs_1st_Order.v:176-199,
 * O/HDL_Source/RL_LJ_Evaluate_Pairs_1st_Order.v:163-176. */
int or_lut_index(float r2, int exp_offset, int nseg);

/* Golden MATLAB pair evaluation, M/Simulation_Verification_Raw_ApoA1_LJ_no_smooth.m:120-198:
 * dx = ref - nbr; r2 = (dx*dx + dy*dy) + dz*dz, unfused; if 0 < r2 < rc2 (strict=1,
 * MATLAB) or 0 < r2 <= rc2 (strict=0, D/RL_LJ_Evaluate_Pairs_1st_Order.v:171-173):
 * t = c1*r2 + c0 (unfused), F = t14 - t8, (Fx,Fy,Fz) = F*(dx,dy,dz); else 0.
 * out: n x 8 floats = F, Fx, Fy, Fz, r2, dx, dy, dz. */
void or_pair_golden(const float *c0_14, const float *c1_14, const float *c0_8,
                    const float *c1_8, int exp_offset, int nseg, float rc2, int strict,
                    const float *ref, const float *nbr, int64_t n, float *out);

/* ---- C model -------------------------------------------------------------- */
typedef struct {
    float eps, sigma, mass, dt, rc;   /* as the reference's literals, R/...Model.cpp:14-24 */
    float box[3];                      /* :26-28 */
    int iterations;                    /* :11 */
} or_cmodel_params;

/* The reference's default constants, computed with the reference's own promotions
 * (R/C_Model_ForceEval/MD_Evaluation_Model.cpp:11-29). */
void or_cmodel_default_params(or_cmodel_params *p);

/* Restatement of MD_Evaluation_Model's iteration loop,
 * R/C_Model_ForceEval/MD_Evaluation_Model.cpp:57-203, on an N x 11 particle array
 * (0-2 pos, 3-5 vel, 6-8 force, 9 LJ energy, 10 kinetic energy, :35).  The array is
 * updated in place.  energies (may be NULL) receives 2 floats per iteration: the
 * printed System_Energy and System_Kinetic_Energy (:158-159).  If `lines` is not NULL,
 * the exact printf text of :159 is appended (caller provides >= 96 bytes per iteration).
 * If `stop_after_force` is set, only the force/kick phase of iteration 0 runs (used by
 * the CPU baseline to time a bounded sample). Returns 1 as the reference does. */
int or_cmodel_run(float *particle, int n, const or_cmodel_params *p, float *energies,
                  char *lines, int stop_after_force);

/* Time-bounded sample of the C model's force loop: evaluates reference atoms
 * [0, nref) against all n atoms (the :63-124 double loop) and returns the number of
 * in-cutoff ordered pairs seen.  Used only as the timed CPU baseline. */
int64_t or_cmodel_force_sample(const float *particle, int n, const or_cmodel_params *p,
                               int nref, float *fout);

/* ---- binning and the half-shell pair stream ------------------------------ */
/* MATLAB binning, M/Simulation_Verification_LJArgon.m:143-175: cell = ceil(p / rc)
 * (1-based, p == 0 -> 1), particles in stable input order.  pos: n x 3 float (already
 * min-shifted).  cell_xyz_out: n x 3 int (1-based; 0 if out of range). */
int or_bin_ceil(const float *pos, int n, float rc, const int ncell[3], int *cell_xyz_out);

/* Regenerates the text of M/VERIFICATION_PARTICLE_PAIR_INPUT.txt for a home cell
 * (M/Simulation_Verification_LJArgon.m:250-572): 8 filters, filter->cell map
 * {0:222, 1:223, 2:231+232, 3:233+311, 4:312+313, 5:321+322, 6:323+331, 7:332+333}
 * relative to the home cell, streams padded to the longest filter with valid bits.
 * raw_xyz: n x 3 double as read from the positions file.  Writes into buf (size cap),
 * returns the number of bytes (or -needed if cap is too small). */
int64_t or_pair_stream_text(const double *raw_xyz, int n, float rc, const int ncell[3],
                            const int home[3], int cell_particle_max, char *buf,
                            int64_t cap);

/* ---- range-limited forces, half-shell, int64 fixed point ----------------- */
enum {
    OR_MODE_PARITY  = 1,   /* unfused arithmetic and dx = (xi - xj) - shift (reference order) */
    OR_MODE_COULOMB = 2,   /* add the r^-3 term */
    OR_MODE_ANALYTIC = 4,  /* FP32 direct r^-14/r^-8/r^-3 instead of the tables, on the
                              correctly rounded separation fl((x_i - t) - e) (t = fl(x_j + S)
                              and e its exact error), the cutoff decided in FP64 */
};

typedef struct {
    int n;
    float box[3];
    int ncell[3];
    float rc2;
    const float *pos4;        /* n x 4: x, y, z (wrapped into [0, L)), scaled charge qs */
    const int *type;          /* n */
    int ntypes;
    const float *pair_a;      /* ntypes^2: table mode 10*eps*sig^12, analytic 480*eps*sig^12 */
    const float *pair_b;      /* ntypes^2: table mode 10*eps*sig^6,  analytic 240*eps*sig^6  */
    const float *tab;         /* interleaved [nseg*256][6] = c0_14 c1_14 c0_8 c1_8 c0_3 c1_3 */
    int exp_offset, nseg;
    int mode;
    const uint32_t *excl;     /* optional exclusion windows, one per atom (NULL = none):
                                 bit d of excl[i] excludes the pair (i, i + d - 16) */
    float rc;                 /* cutoff; analytic mode decides r^2 <= rc*rc exactly (FP64) */
} or_rl_system;

/* Cell index of one coordinate: (int)(x * inv_edge) clamped to [0, nc-1], with
 * inv_edge = (float)(nc / L).  Shared rule of the oracle and the GPU binning. */
int or_cell_of(float x, float inv_edge, int nc);

/* Per-pair fixed-point conversion: rint(f * 2^32) saturated at +-2^62. */
int64_t or_fix(float f);

/* OpenMP threads of the calling thread's later oracle calls (n >= 1); callers that run many
 * small systems side by side give each of their own threads one.  Integer sums make every
 * result independent of the thread count. */
void or_set_threads(int n);

/* Half-shell forces: cells of edge L/ncell, home cell + the 13 offsets of
 * D/Particle_Pair_Gen_HalfShell_64_Cells.v:213-226 ({223,231,232,233,311,312,313,
 * 321,322,323,331,332,333} around 222), explicit image shifts, filter 0 < r2 <= rc2
 * (D/Filter_Logic.v:234), table force (D/RL_LJ_Evaluate_Pairs_1st_Order.v:409-475,
 * v1 per-pair multipliers D/RL_Evaluate_Pairs_1st_Order.v:53-55, :325-390),
 * reference particle +F and neighbour -F (D/RL_LJ_Evaluation_Unit.v:137), each pair
 * once (N3L).  Every pair contribution is converted to int64 (or_fix) before it is
 * added, so the result is independent of summation order.
 * fix_out: n x 3 int64 (2^-32 kJ/mol/nm units).  Returns the number of pair-evals. */
int64_t or_rl_forces_fix(const or_rl_system *s, int64_t *fix_out);

/* FP64 analytic oracle: LJ (Lorentz-Berthelot params given as eps/sig per type pair)
 * plus cutoff Coulomb, all periodic images within rc, forces in kJ/mol/A.
 * qs: plain charges (e).  Brute force over all pairs and the 27 nearest images, or the
 * cell list when use_cells is set.  Optional energy (kJ/mol) and pair virial
 * W[3a+b] = sum over pairs of F d_a d_b (F d = force on i, d = r_i - r_j; 9 values). */
int64_t or_fp64_direct(int n, const double *xyz, const int *type, const double *q,
                       int ntypes, const double *eps_ij, const double *sig_ij,
                       double ke, int coulomb, const double box[3], double rc,
                       int use_cells, double *f_out, double *energy, double *virial,
                       const uint32_t *excl);

/* FP64 evaluation of the same first-order table model (error measure (i)). */
int64_t or_table_fp64(const or_rl_system *s, double *f_out);

/* ---- Motion_Update ---------------------------------------------------------- */
/* Fused velocity-Verlet update on the int64 forces (the GPU's k_motion_update):
 *   dv  = (float)((double)fix * hk[type])      hk = 2^-32 * 1e-5 / m * dt/2
 *   v   = vh + dv   (full-step velocity; skipped on the first step: v = vh)
 *   vh' = v + dv;  x' = x + vh'*dt; wrap into [0, L)
 * kick_drift=1 gives the reference's kick-then-drift (C model :133-150, :164-202;
 * M/FullScale_Simulation_LJArgon.m:1062-1074): v' = v + 2*dv; x' = x + v'*dt.
 * ke_out (optional) receives sum 0.5 m v^2 in kJ/mol using the full-step v. */
void or_motion_update(int n, float *pos4, float *vel3, const int64_t *fix,
                      const int *type, const double *hk, const double *mass, float dt,
                      const float box[3], int first, int kick_drift, double *ke_out);

/* Capped steepest descent step (libmdrl's untimed pre-relaxation):
 * F = fix * 2^-32, s = min(gamma, dmax / |F|), x' = wrap(x + (float)(F * s)). */
void or_minimize_update(int n, float *pos4, const int64_t *fix, double gamma, double dmax,
                        const float box[3]);

/* ---- bonded terms (the bonded-chain workload) ------------------------------------ */
/* One term per row of atoms[nterms][4] (input indices, -1 pads): a bond (i, j, -1, -1),
 * an angle (i, j, k, -1) with vertex j, or a torsion (i, j, k, l).  params[nterms][4],
 * forces in kJ/mol/nm from positions in A (the constants carry the factor 10):
 *   bond    (r0, 10 k, -, -):          U = k (r - r0)^2        Bonded/calc_bond_v2.v:83-110
 *   angle   (theta0, 10 k, -, -):      U = k (theta - theta0)^2  Bonded/calc_angle_v2.v:113-147
 *   torsion (10 k, n, cos d, sin d):   U = k (1 + cos(n phi + d))  Bonded/calcTorsion_v2.v:70-274
 * Vectors between term atoms take the minimum image of the box.  theta uses or_atan2f
 * (Taylor series of arctan after a reduction by pi/6); phi enters only through
 * cos(n phi) and sin(n phi), built by the angle-addition recurrence from cos phi and
 * sin phi.  Torsion forces follow the projection form of Bekker (1996) / Blondel and
 * Karplus (1996), with no division by sin phi. */
float or_atan2f(float y, float x);
/* FP32 forces, each atom's share converted by or_fix and added into fix_out (n x 3,
 * accumulated onto its contents).  term_energy (may be NULL): FP64 energy per term from
 * the same FP32 positions, in kJ/mol (k taken as params / 10). */
void or_bonded_fix(int n, const float *pos4, const float box[3], int nterms,
                   const int32_t *atoms, const float *params, int64_t *fix_out,
                   double *term_energy);
/* The same potentials in FP64 (forces in kJ/mol/A, k in kJ/mol/A^2 etc. without the
 * factor 10, torsion phase in params[2] as an angle): forces added into f_out, total
 * energy returned.  Used to check the FP32 version and, by finite differences, itself. */
double or_bonded_fp64(int n, const double *xyz, const double box[3], int nterms,
                      const int32_t *atoms, const double *params, double *f_out);

#ifdef __cplusplus
}
#endif
#endif
