/*
 * mcl_driver — the reference's driver loop (Main.c:3-79) over libmcl_hip.so.
 *
 *   mcl_driver <input file>
 *
 * Reads the reference's input format, builds the whole mesh on the host
 * (quadtree, connectivity, boundary conditions), uploads it once, sets the
 * initial temperature, and runs the per-step sequence of Main.c:58-64 on the
 * device.  Prints the reference's progress lines ("%d dv = ... dp = ..." per
 * Stokes solve, "Temperature CG cycle = %d" per heat step) and writes the
 * reference's output files every save_spacing steps.  Exits non-zero on any
 * error (the reference exits 0 from terminate(), Parallel.c:345-352).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mcl.h"
#include "mcl_host.h"

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int die(mcl_ctx *ctx, const char *what, int rc)
{
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, ctx ? mcl_last_error(ctx) : "");
    return 1;
}

int main(int argc, char **argv)
{
    mclh_input in;
    mclh_mesh *mesh = NULL;
    mcl_ctx *ctx = NULL;
    mcl_mesh_view views[32];
    mcl_control ctl;
    char err[512];
    int rc, k, step;
    double t0;
    const mclh_level *fin;
    double *T = NULL, *S = NULL, *Vis = NULL, *V = NULL;

    if (argc < 2) {
        fprintf(stderr, "Usage: %s inputfile\n", argv[0]);
        return 10;
    }
    if (mclh_read_input(argv[1], &in, err, sizeof err)) {
        fprintf(stderr, "%s\n", err);
        return 11;
    }
    t0 = now();
    if (mclh_build_mesh(&in, &mesh, err, sizeof err)) {
        fprintf(stderr, "%s\n", err);
        return 12;
    }
    fprintf(stderr, "Tree setup done!\n");
    if ((rc = mcl_create(in.device, &ctx))) return die(NULL, "mcl_create", rc);
    for (k = 0; k < mesh->nlev; k++) mclh_fill_view(mesh, k, &views[k]);
    if ((rc = mcl_mesh_upload(ctx, views, mesh->nlev))) return die(ctx, "mcl_mesh_upload", rc);
    mclh_fill_control(&in, &ctl);
    if ((rc = mcl_set_control(ctx, &ctl))) return die(ctx, "mcl_set_control", rc);
    fin = &mesh->lev[mesh->nlev - 1];
    T = malloc(fin->nno * sizeof(double));
    S = malloc(fin->nno * sizeof(double));
    Vis = malloc(fin->nno * sizeof(double));
    V = malloc(2 * fin->nno * sizeof(double));
    for (k = 0; k < mesh->nlev; k++) {
        double *Tk = malloc(mesh->lev[k].nno * sizeof(double));
        if (mclh_initial_temperature(&in, mesh, k, Tk, err, sizeof err)) {
            fprintf(stderr, "%s\n", err);
            return 13;
        }
        if ((rc = mcl_set_field(ctx, mesh->lev[k].lev, MCL_FIELD_T, Tk, mesh->lev[k].nno)))
            return die(ctx, "mcl_set_field", rc);
        free(Tk);
    }
    mcl_set_model_time(ctx, 0.0);
    for (step = 0; step <= in.max_timestep; step++) {
        mcl_stokes_stats ss;
        mcl_heat_stats hs;
        double mt = 0.0;
        mcl_get_model_time(ctx, &mt);
        fprintf(stderr, "\ntimestep = %d, CPU time = %.0fs, model time = %e\n", step, now() - t0, mt);
        if ((rc = mcl_buoyancy_force(ctx))) return die(ctx, "mcl_buoyancy_force", rc);
        if ((rc = mcl_viscosity(ctx))) return die(ctx, "mcl_viscosity", rc);
        if ((rc = mcl_velocity_pressure_KGF_matrix(ctx))) return die(ctx, "mcl_velocity_pressure_KGF_matrix", rc);
        fprintf(stderr, "citcom solver\n");
        if ((rc = mcl_solve_velocity_pressure(ctx, &ss))) return die(ctx, "mcl_solve_velocity_pressure", rc);
        printf("%d dv = %.10e dp = %.10e\n", ss.uzawa_cycles, ss.dv, ss.dp);
        if (in.write_output && in.save_spacing > 0 && step % in.save_spacing == 0) {
            mcl_get_field(ctx, fin->lev, MCL_FIELD_T, T, fin->nno);
            mcl_get_field(ctx, fin->lev, MCL_FIELD_STRESS, S, fin->nno);
            mcl_get_field(ctx, fin->lev, MCL_FIELD_VIS, Vis, fin->nno);
            mcl_get_field(ctx, fin->lev, MCL_FIELD_V, V, 2 * (size_t)fin->nno);
            if (mclh_write_output(&in, fin, step, T, S, Vis, V, err, sizeof err)) {
                fprintf(stderr, "%s\n", err);
                return 14;
            }
        }
        if ((rc = mcl_update_temperature(ctx, &hs))) return die(ctx, "mcl_update_temperature", rc);
        /* the first step's initial_T_dot solve prints its own line first
         * (Update_temperature.c:1128, 1255) */
        if (step == 0) printf("Temperature CG cycle = %d\n", hs.init_cg_cycles);
        printf("Temperature CG cycle = %d\n", hs.cg_cycles);
        fflush(stdout);
    }
    mcl_get_field(ctx, fin->lev, MCL_FIELD_T, T, fin->nno);
    mcl_get_field(ctx, fin->lev, MCL_FIELD_V, V, 2 * (size_t)fin->nno);
    {
        double vrms, nu;
        mclh_diagnostics(&in, fin, T, V, &vrms, &nu);
        fprintf(stderr, "Vrms = %.10e Nu_top = %.10e\n", vrms, nu);
    }
    fprintf(stderr, "%f seconds used in total\n", now() - t0);
    free(T); free(S); free(Vis); free(V);
    mcl_destroy(ctx);
    mclh_free_mesh(mesh);
    return 0;
}
