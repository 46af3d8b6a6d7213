/*
 * Golden-vector harness for the reference solver (test infrastructure only).
 *
 * It is linked against the reference's own object files, compiled from the
 * unmodified sources, in place of the reference's Main.o.  It runs the same
 * per-step sequence as the reference driver (Main.c:28-69) and, after each
 * phase, appends the reference's global state G to a binary dump so that the
 * build's restatement and HIP path can be compared with it.
 *
 * Dump format (little endian), one record per array:
 *   char   name[48]   (NUL padded)
 *   char   kind       ('i' int32, 'd' float64, 'c' int8)
 *   char   pad[7]
 *   int64  count
 *   <count elements>
 *
 * Usage: mcl_ref_harness <input file> <dump file> [full_steps] [snapshot_every]
 *   full_steps      steps (from step 0) whose matrices and vectors are dumped
 *   snapshot_every  every k-th step dumps T, V, p (0 = never)
 * Every step dumps dt and model_time.
 */
#include "global_variables.h"

static FILE *dump_fp;

static void rec(const char *name, char kind, const void *data, long long count)
{
    char hdr[56];
    size_t esz = kind == 'd' ? 8 : (kind == 'i' ? 4 : 1);
    memset(hdr, 0, sizeof(hdr));
    strncpy(hdr, name, 47);
    hdr[48] = kind;
    fwrite(hdr, 1, 56, dump_fp);
    fwrite(&count, sizeof(long long), 1, dump_fp);
    if (count > 0)
        fwrite(data, esz, (size_t)count, dump_fp);
}

static void rec_i1(const char *name, int v) { rec(name, 'i', &v, 1); }
static void rec_d1(const char *name, double v) { rec(name, 'd', &v, 1); }

/* Row-slot matrix (ncol[row], col[row][slot], val[row][slot]) as CSR. */
static void rec_rows(const char *prefix, int nrow, int *ncol, int **col, double **val)
{
    char name[64];
    int i, j, nnz = 0, k = 0;
    int *ptr = malloc((nrow + 1) * sizeof(int));
    int *c;
    double *v;
    ptr[0] = 0;
    for (i = 0; i < nrow; i++) { nnz += ncol[i]; ptr[i + 1] = nnz; }
    c = malloc((nnz + 1) * sizeof(int));
    v = malloc((nnz + 1) * sizeof(double));
    for (i = 0; i < nrow; i++)
        for (j = 0; j < ncol[i]; j++) { c[k] = col[i][j]; v[k] = val[i][j]; k++; }
    snprintf(name, sizeof(name), "%s.ptr", prefix); rec(name, 'i', ptr, nrow + 1);
    snprintf(name, sizeof(name), "%s.col", prefix); rec(name, 'i', c, nnz);
    snprintf(name, sizeof(name), "%s.val", prefix); rec(name, 'd', v, nnz);
    free(ptr); free(c); free(v);
}

static void dump_setup(void)
{
    char name[64];
    int lev, a, d;
    rec_i1("max_level", G.max_level);
    rec_i1("min_level", G.min_level);
    rec_i1("gnox", G.gnox);
    rec_i1("gnoz", G.gnoz);
    rec("dx", 'd', G.dx, G.max_level);
    rec("dz", 'd', G.dz, G.max_level);
    for (lev = G.min_level; lev < G.max_level; lev++) {
#define NM(fmt) (snprintf(name, sizeof(name), fmt, lev), name)
        rec_i1(NM("L%d.nel"), G.nel[lev]);
        rec_i1(NM("L%d.nno"), G.nno[lev]);
        rec_i1(NM("L%d.neq"), G.neq[lev]);
        rec_i1(NM("L%d.heat.neq"), G.heat[lev].neq);
        rec(NM("L%d.Mcode"), 'i', G.Mcode[lev], G.nel[lev]);
        for (a = 0; a < NEN; a++) {
            snprintf(name, sizeof(name), "L%d.ien%d", lev, a);
            rec(name, 'i', G.ien[lev][a], G.nel[lev]);
        }
        rec(NM("L%d.node"), 'i', G.node[lev], G.nno[lev]);
        for (d = 0; d < NSD; d++) {
            snprintf(name, sizeof(name), "L%d.id%d", lev, d);
            rec(name, 'i', G.id[lev][d], G.nno[lev]);
            snprintf(name, sizeof(name), "L%d.v_bc%d", lev, d);
            rec(name, 'c', G.bc.v_bc[lev][d], G.nno[lev]);
            snprintf(name, sizeof(name), "L%d.v_bc_val%d", lev, d);
            rec(name, 'd', G.bc.v_bc_val[lev][d], G.nno[lev]);
            snprintf(name, sizeof(name), "L%d.X%d", lev, d);
            rec(name, 'd', G.X[lev][d], G.nno[lev]);
            for (a = 0; a < NEN; a++) {
                snprintf(name, sizeof(name), "L%d.lm%d_%d", lev, d, a);
                rec(name, 'i', G.lm[lev][d][a], G.nel[lev]);
            }
        }
        rec(NM("L%d.veq_node"), 'i', G.veq_node[lev], G.neq[lev]);
        rec(NM("L%d.veq_dof"), 'i', G.veq_dof[lev], G.neq[lev]);
        rec(NM("L%d.T_bc"), 'c', G.bc.T_bc[lev], G.nno[lev]);
        rec(NM("L%d.T_bc_val"), 'd', G.bc.T_bc_val[lev], G.nno[lev]);
        rec(NM("L%d.heat.id"), 'i', G.heat[lev].id, G.nno[lev]);
        rec(NM("L%d.heat.eq_node"), 'i', G.heat[lev].eq_node, G.heat[lev].neq);
        for (a = 0; a < NEN; a++) {
            snprintf(name, sizeof(name), "L%d.heat.lm%d", lev, a);
            rec(name, 'i', G.heat[lev].lm[a], G.nel[lev]);
        }
        rec_i1(NM("L%d.hangle.nh"), G.hangle[lev].nh);
        rec(NM("L%d.hangle.node"), 'i', G.hangle[lev].node, G.hangle[lev].nh);
        rec(NM("L%d.hangle.ni"), 'i', G.hangle[lev].ni, G.hangle[lev].nh);
        for (d = 0; d < NSD; d++) {
            snprintf(name, sizeof(name), "L%d.hangle.cnode%d", lev, d);
            rec(name, 'i', G.hangle[lev].cnode[d], G.hangle[lev].nh);
            snprintf(name, sizeof(name), "L%d.hangle.ci%d", lev, d);
            rec(name, 'i', G.hangle[lev].ci[d], G.hangle[lev].nh);
        }
        rec(NM("L%d.dx_dxi"), 'd', G.dx_dxi[lev], G.nel[lev]);
        rec(NM("L%d.dz_deta"), 'd', G.dz_deta[lev], G.nel[lev]);
        rec(NM("L%d.jacobi"), 'd', G.jacobi[lev], G.nel[lev]);
        rec(NM("L%d.area"), 'd', G.eco[lev].area, G.nel[lev]);
        rec(NM("L%d.p2n.nel"), 'i', G.p2n[lev].nel, G.nno[lev]);
        for (a = 0; a < NEN; a++) {
            snprintf(name, sizeof(name), "L%d.p2n.el%d", lev, a);
            rec(name, 'i', G.p2n[lev].el[a], G.nno[lev]);
        }
        if (strcmp(G.solver, "mg") == 0 && lev < G.max_level - 1) {
            int lv = lev + 1, m = G.neq[lv], k;
            int *tmp;
            /* interp[upper level]: fine eq <- average of local_eq_n coarse eqs */
            rec(NM("L%d.interp.local_eq_n"), 'i', G.interp[lv].local_eq_n, m);
            tmp = malloc(4 * (size_t)m * sizeof(int));
            for (k = 0; k < 4 * m; k++) tmp[k] = -7;
            for (a = 0; a < 4; a++)
                for (k = 0; k < m; k++)
                    if (G.interp[lv].local_eq_n[k] > a)
                        tmp[a * m + k] = G.interp[lv].local_eq[a][k];
            rec(NM("L%d.interp.local_eq"), 'i', tmp, 4 * (long long)m);
            free(tmp);
            /* project[lower level]: injection pairs (local_l, local_u) */
            rec_i1(NM("L%d.project.nlocal"), G.project[lev].nlocal);
            rec(NM("L%d.project.local_l"), 'i', G.project[lev].local_l, G.project[lev].nlocal);
            rec(NM("L%d.project.local_u"), 'i', G.project[lev].local_u, G.project[lev].nlocal);
        }
#undef NM
    }
}

static void dump_fields(const char *tag, int step)
{
    char name[96];
    int lev;
    for (lev = G.min_level; lev < G.max_level; lev++) {
        snprintf(name, sizeof(name), "s%d.L%d.T.%s", step, lev, tag);
        rec(name, 'd', G.T[lev], G.nno[lev]);
    }
}

static void dump_assembly(int step)
{
    char name[96];
    int lev;
    for (lev = G.min_level; lev < G.max_level; lev++) {
        int neq = G.neq[lev];
        snprintf(name, sizeof(name), "s%d.L%d.buoyancy", step, lev);
        rec(name, 'd', G.buoyancy[lev], G.nno[lev]);
        snprintf(name, sizeof(name), "s%d.L%d.vis", step, lev);
        rec(name, 'd', G.vis[lev], G.nno[lev]);
        snprintf(name, sizeof(name), "s%d.L%d.K", step, lev);
        rec_rows(name, neq, G.Kncol[lev], G.Krow_col[lev], G.Krow_val[lev]);
        snprintf(name, sizeof(name), "s%d.L%d.G", step, lev);
        rec_rows(name, neq, G.Gncol[lev], G.Grow_col[lev], G.Grow_val[lev]);
        snprintf(name, sizeof(name), "s%d.L%d.Kdiag", step, lev);
        rec(name, 'd', G.Kdiag[lev], neq);
        snprintf(name, sizeof(name), "s%d.L%d.F", step, lev);
        rec(name, 'd', G.F[lev], neq);
    }
}

static void dump_solution(int step)
{
    char name[96];
    int lev = G.max_level - 1;
    snprintf(name, sizeof(name), "s%d.d", step); rec(name, 'd', G.d[lev], G.neq[lev]);
    snprintf(name, sizeof(name), "s%d.p", step); rec(name, 'd', G.p[lev], G.nel[lev]);
    snprintf(name, sizeof(name), "s%d.V0", step); rec(name, 'd', G.V[lev][0], G.nno[lev]);
    snprintf(name, sizeof(name), "s%d.V1", step); rec(name, 'd', G.V[lev][1], G.nno[lev]);
    snprintf(name, sizeof(name), "s%d.stress", step); rec(name, 'd', G.stress[lev], G.nno[lev]);
}

static void dump_heat(int step)
{
    char name[96];
    int lev = G.max_level - 1;
    struct HEAT *h = &G.heat[lev];
    snprintf(name, sizeof(name), "s%d.heat.tau0", step); rec(name, 'd', h->tau[0], G.nel[lev]);
    snprintf(name, sizeof(name), "s%d.heat.tau1", step); rec(name, 'd', h->tau[1], G.nel[lev]);
    snprintf(name, sizeof(name), "s%d.heat.K", step);
    rec_rows(name, h->neq, h->Kncol, h->Krow_col, h->Krow_val);
    snprintf(name, sizeof(name), "s%d.heat.C", step);
    rec_rows(name, h->neq, h->Cncol, h->Crow_col, h->Crow_val);
    snprintf(name, sizeof(name), "s%d.heat.F", step); rec(name, 'd', h->F, h->neq);
    snprintf(name, sizeof(name), "s%d.heat.d", step); rec(name, 'd', h->d, h->neq);
    snprintf(name, sizeof(name), "s%d.heat.ddot", step); rec(name, 'd', h->ddot, h->neq);
    snprintf(name, sizeof(name), "s%d.T", step); rec(name, 'd', G.T[lev], G.nno[lev]);
}

int main(int argc, char **argv)
{
    int full_steps = 1, snap_every = 0, write_files = 0;
    char name[96];
    int lev;
    if (argc < 3) {
        fprintf(stderr, "usage: %s input dump [full_steps] [snapshot_every] [write_output]\n", argv[0]);
        return 2;
    }
    if (argc > 3) full_steps = atoi(argv[3]);
    if (argc > 4) snap_every = atoi(argv[4]);
    if (argc > 5) write_files = atoi(argv[5]);

    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &G.ip);
    MPI_Comm_size(MPI_COMM_WORLD, &G.np);
    {
        /* the reference parser reads argv[1] */
        char *av[2];
        av[0] = argv[0];
        av[1] = argv[1];
        setup_parser(2, av);
    }
    read_input_file();
    if (G.np != G.npx * G.npz) {
        fprintf(stderr, "rank count mismatch\n");
        MPI_Finalize();
        return 3;
    }
    dump_fp = fopen(argv[2], "wb");
    if (!dump_fp) { perror("dump"); return 4; }

    initial_data_structure();
    get_time_independent_variables();

    G.timestep = 0;
    G.solution_cycle = 0;
    while (G.timestep <= G.max_timestep) {
        int full = G.timestep < full_steps;
        if (G.solution_cycle == 0 || G.control.AMR == 1) {
            build_quad_tree();
            distribute_Mcode_among_processors();
            construct_ien();
            node_index_from_ien();
            node_position();
            element_area();
            what_elements_node_in();
            boundary_condition();
            construct_id();
            construct_lm();
            construct_connection();
            construct_shape_function();
            assign_matrix_memory();
            if (G.solution_cycle == 0)
                dump_setup();
        }
        if (G.solution_cycle == 0) {
            initial_temperature();
            dump_fields("init", 0);
        }
        buoyancy_force();
        viscosity();
        velocity_pressure_KGF_matrix();
        if (full) dump_assembly(G.timestep);
        fflush(stdout);
        solve_velocity_pressure();
        fflush(stdout);
        if (full) dump_solution(G.timestep);
        /* the reference's own output files, Main.c:62-63 */
        if (write_files && G.timestep % G.save_spacing == 0)
            output();
        update_temperature();
        fflush(stdout);
        if (full) dump_heat(G.timestep);
        lev = G.max_level - 1;
        snprintf(name, sizeof(name), "s%d.dt", G.timestep); rec_d1(name, G.heat[lev].dt);
        snprintf(name, sizeof(name), "s%d.model_time", G.timestep); rec_d1(name, G.model_time);
        if (snap_every > 0 && !full && (G.timestep % snap_every == 0 || G.timestep == G.max_timestep)) {
            snprintf(name, sizeof(name), "s%d.T", G.timestep); rec(name, 'd', G.T[lev], G.nno[lev]);
            snprintf(name, sizeof(name), "s%d.V0", G.timestep); rec(name, 'd', G.V[lev][0], G.nno[lev]);
            snprintf(name, sizeof(name), "s%d.V1", G.timestep); rec(name, 'd', G.V[lev][1], G.nno[lev]);
            snprintf(name, sizeof(name), "s%d.p", G.timestep); rec(name, 'd', G.p[lev], G.nel[lev]);
        }
        fflush(dump_fp);
        if (G.control.AMR == 1)
            release_memory();
        G.timestep++;
        G.solution_cycle++;
    }
    rec_i1("steps_done", G.timestep);
    fclose(dump_fp);
    MPI_Finalize();
    return 0;
}
