/*
 * host_c.c -- a plain C host linking libmdrl in place of the reference's C model
 * (Range_Limited/C_Model_ForceEval/MD_Evaluation_Model.h:1 declares the one entry point
 * the reference offers; its only caller, Range_Limited/Module_Gen_Scripts/main.cpp:110,
 * is commented out).
 *
 *   host_c cmodel          calls MD_Evaluation_Model() exactly as a reference host would
 *                          (input PDB from $MDRL_INPUT); exit status 0 when it returns 1
 *   host_c steps N [out]   the 512-atom argon system through the half-shell system path:
 *                          N velocity-Verlet steps, then the int64 forces (N x 3, input
 *                          order) written to `out` as raw little-endian int64 and one
 *                          summary line on stdout
 *
 * Build: cc host_c.c -I include -L sc2019_md_amd -lmdrl -Wl,-rpath,<dir of libmdrl.so>
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdrl.h"

static int fail(const char *what, int rc)
{
    fprintf(stderr, "host_c: %s failed (%d): %s\n", what, rc, md_rl_last_error());
    return 1;
}

static int run_steps(int nsteps, const char *out)
{
    md_rl_system s;
    int rc = md_rl_make_system(&s, 1, 0, 2019);
    if (rc) return fail("md_rl_make_system", rc);
    md_rl_params p;
    memset(&p, 0, sizeof p);
    p.n = s.n;
    for (int d = 0; d < 3; ++d) p.box[d] = s.box[d];
    p.rc = s.rc;
    p.dt = s.dt;
    p.ntypes = s.ntypes;
    p.eps = s.eps;
    p.sigma = s.sigma;
    p.mass = s.mass;
    p.coulomb_k = s.coulomb_k;
    md_rl_ctx *ctx = NULL;
    if ((rc = md_rl_create(&ctx, &p))) { md_rl_free_system(&s); return fail("md_rl_create", rc); }
    int64_t *fix = (int64_t *)malloc(sizeof(int64_t) * 3 * (size_t)s.n);
    double pe = 0.0, ke = 0.0, vir[9];
    md_rl_stats st;
    if (!fix) { rc = MDRL_ENOMEM; goto done; }
    if ((rc = md_rl_set_types(ctx, s.type, s.charge, s.n)) ||
        (rc = md_rl_set_particles(ctx, s.particle, s.n)) ||
        (rc = md_rl_step(ctx, nsteps)) ||
        (rc = md_rl_get_forces_fixed(ctx, fix, s.n)) ||
        (rc = md_rl_energy(ctx, &pe, &ke, vir)) ||
        (rc = md_rl_get_stats(ctx, &st)))
        goto done;
    if (out) {
        FILE *f = fopen(out, "wb");
        if (!f || fwrite(fix, sizeof(int64_t), 3 * (size_t)s.n, f) != 3 * (size_t)s.n) {
            if (f) fclose(f);
            fprintf(stderr, "host_c: cannot write %s\n", out);
            rc = MDRL_EIO;
            goto done;
        }
        fclose(f);
    }
    printf("atoms %d steps %d pairs %lld pe %.9e ke %.9e\n", s.n, nsteps, (long long)st.pair_evals, pe, ke);
done:
    free(fix);
    md_rl_destroy(ctx);
    md_rl_free_system(&s);
    return rc ? fail("system path", rc) : 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && strcmp(argv[1], "cmodel") == 0) return MD_Evaluation_Model() == 1 ? 0 : 1;
    if (argc >= 3 && strcmp(argv[1], "steps") == 0) return run_steps(atoi(argv[2]), argc >= 4 ? argv[3] : NULL);
    fprintf(stderr, "usage: %s cmodel | steps N [out]\n", argv[0]);
    return 2;
}
