/*
 * mcl_host.c — C host layer: input parser, quadtree mesh, initial temperature,
 * output files and diagnostics.  See mcl_host.h.
 *
 * Integer results (Mcode, ien, node, id, lm, BC flags, hanging nodes, MG maps)
 * follow the reference's rules exactly; they are compared element for element
 * with the reference's own arrays in the tests.
 */
#define _GNU_SOURCE
#include "mcl_host.h"

#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

static int herr(char *err, int errlen, const char *fmt, ...)
{
    if (err && errlen > 0) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, errlen, fmt, ap);
        va_end(ap);
    }
    return 1;
}

/* ------------------------------------------------------------------ input */

void mclh_default_input(mclh_input *in)
{
    /* defaults of Input.c:3-98 */
    memset(in, 0, sizeof *in);
    strcpy(in->data_dir, ".");
    strcpy(in->solver, "jc");
    in->max_level = 1;
    in->min_level = 1;
    in->npx = 2;
    in->npz = 2;
    in->length = 2.0;
    in->height = 1.0;
    memcpy(in->top_v_bc, "sv", 2);
    memcpy(in->bot_v_bc, "sv", 2);
    memcpy(in->lef_v_bc, "vs", 2);
    memcpy(in->rig_v_bc, "vs", 2);
    in->top_T_bc = 'T';
    in->bot_T_bc = 'T';
    in->lef_T_bc = 'F';
    in->rig_T_bc = 'F';
    in->top_T_bc_val = 0.0;
    in->bot_T_bc_val = 1.0;
    in->T_ini = 1;
    in->T_mantle = 0.5;
    in->T_mantle_perturb = 0.01;
    in->T_wave_number = 6;
    in->Ra = 1e7;
    in->refT = 0.5;
    in->lm_jump = 1.0;
    in->highest_vis = 1.0;
    in->lowest_vis = 1.0;
    in->use_dense_matrix = 1;
    in->mg_coarse_iteration = 4;
    in->mg_finest_iteration = 4;
    in->save_spacing = 1;
    in->mpi_KF_method = 2;
    in->accuracy = 0.001;
    in->citcom_accu = 0.001;
    in->dev_mode = 1;
    in->max_citcom_cycle = 375;
    in->mg_cycle = 2;
    in->max_cg_cycle = 200;
    in->old_T_level = 1;
    in->zero_P = 1;
    /* keys of this build */
    in->inner_rtol = 3e-7;
    in->T_seed = 20261202ULL;
    in->write_output = 1;
}

static void copy_str(char *dst, size_t cap, const char *src)
{
    snprintf(dst, cap, "%s", src);
}

/* One key=value (values already unquoted).  Numbers are read like the
 * reference's sscanf("%d") / sscanf("%lf").  Returns 0 if the key is known. */
int mclh_set_key(mclh_input *in, const char *k, const char *v)
{
#define I(name, field) if (!strcmp(k, name)) { sscanf(v, "%d", &in->field); return 0; }
#define D(name, field) if (!strcmp(k, name)) { sscanf(v, "%lf", &in->field); return 0; }
#define C1(name, field) if (!strcmp(k, name)) { in->field = v[0]; return 0; }
#define C2(name, field) if (!strcmp(k, name)) { in->field[0] = v[0]; in->field[1] = v[0] ? v[1] : 0; return 0; }
#define D2(name, field) if (!strcmp(k, name)) { sscanf(v, "%lf", &in->field[0]); sscanf(v, "%*f,%lf", &in->field[1]); return 0; }
    if (!strcmp(k, "data_dir")) { copy_str(in->data_dir, sizeof in->data_dir, v); return 0; }
    if (!strcmp(k, "solver")) { copy_str(in->solver, sizeof in->solver, v); return 0; }
    if (!strcmp(k, "old_T_file")) { copy_str(in->old_T_file, sizeof in->old_T_file, v); return 0; }
    I("level", max_level) I("mg_min_level", min_level) I("nprox", npx) I("nproz", npz)
    D("model_length", length) D("model_height", height)
    C2("top_v_bc", top_v_bc) C2("bot_v_bc", bot_v_bc) C2("lef_v_bc", lef_v_bc) C2("rig_v_bc", rig_v_bc)
    D2("top_v_bc_val", top_v_bc_val) D2("bot_v_bc_val", bot_v_bc_val)
    D2("lef_v_bc_val", lef_v_bc_val) D2("rig_v_bc_val", rig_v_bc_val)
    C1("top_T_bc", top_T_bc) C1("bot_T_bc", bot_T_bc) C1("lef_T_bc", lef_T_bc) C1("rig_T_bc", rig_T_bc)
    D("top_T_bc_val", top_T_bc_val) D("bot_T_bc_val", bot_T_bc_val)
    D("lef_T_bc_val", lef_T_bc_val) D("rig_T_bc_val", rig_T_bc_val)
    I("T_ini", T_ini) D("T_mantle", T_mantle) D("T_mantle_perturb", T_mantle_perturb)
    I("T_wave_number", T_wave_number) D("Rayleigh", Ra) I("use_hash", use_hash)
    D("A", A) D("refT", refT) D("lower_mantle_jump", lm_jump) D("highest_vis", highest_vis)
    D("lowest_vis", lowest_vis) I("use_dense_matrix", use_dense_matrix)
    I("mg_coarse_iteration", mg_coarse_iteration) I("mg_finest_iteration", mg_finest_iteration)
    I("max_timestep", max_timestep) I("save_spacing", save_spacing) D("Q", Q) D("theta", theta)
    I("AMR", AMR) I("mpi_KF_method", mpi_KF_method) I("cg_precondition", cg_precondition)
    I("Tcg_precondition", Tcg_precondition) D("accuracy", accuracy) I("auto_accuracy", auto_accuracy)
    D("citcom_solver_accu", citcom_accu) I("develop_mode", dev_mode) I("check_v_only", check_v_only)
    I("max_citcom_cycle", max_citcom_cycle) I("mg_cycle", mg_cycle) I("max_cg_cycle", max_cg_cycle)
    I("old_T_level", old_T_level) I("old_T_timestep", old_T_timestep) I("lg_mul", lg_mul)
    I("grain_damage", grain_damage) I("show_convg", show_convg) I("zero_P", zero_P)
    /* keys of this build */
    I("schur_mode", schur) I("precision", precision) I("warm_start", warm_start) D("inner_rtol", inner_rtol)
    I("exact_dots", exact_dots)
    I("device", device) I("write_output", write_output)
    if (!strcmp(k, "T_seed")) { in->T_seed = strtoull(v, NULL, 10); return 0; }
#undef I
#undef D
#undef C1
#undef C2
#undef D2
    return 1;
}

/* Parsing.c:54-139 (setup_parser): entries name=value separated by blanks,
 * a value may be quoted with " or ', '#' starts a comment, and a later entry
 * of the same name replaces an earlier one. */
int mclh_read_input(const char *path, mclh_input *in, char *err, int errlen)
{
    FILE *fp = fopen(path, "r");
    char line[1024];
    mclh_default_input(in);
    if (!fp) return herr(err, errlen, "File: %s is unreadable", path);
    while (fgets(line, sizeof line, fp)) {
        char *pl = line;
        for (;;) {
            char name[64], value[1024], *pn = name, *pv = value, t1, t2;
            while (*pl == ' ' || *pl == '\t') pl++;
            if (*pl == '\0' || *pl == '\n' || *pl == '#') break;
            while (*pl != '=' && *pl != '\0' && *pl != ' ' && *pl != '\n' && *pl != '\t' &&
                   pn < name + sizeof name - 1)
                *pn++ = *pl++;
            *pn = '\0';
            if (*pl == '=') pl++;
            if (*pl == '"' || *pl == '\'') {
                t1 = t2 = *pl++;
            } else {
                t1 = ' ';
                t2 = '\t';
            }
            while (*pl != t1 && *pl != t2 && *pl != '\0' && *pl != '\n' && pv < value + sizeof value - 1)
                *pv++ = *pl++;
            *pv = '\0';
            if (*pl == '"' || *pl == '\'') pl++;
            mclh_set_key(in, name, value);
        }
    }
    fclose(fp);
    if (in->min_level >= in->max_level)
        return herr(err, errlen, "min_level must be smaller than max_level");
    return 0;
}

void mclh_fill_control(const mclh_input *in, mcl_control *c)
{
    memset(c, 0, sizeof *c);
    c->Ra = in->Ra;
    c->A = in->A;
    c->refT = in->refT;
    c->lm_jump = in->lm_jump;
    c->accuracy = in->accuracy;
    c->citcom_accu = in->citcom_accu;
    c->theta = in->theta;
    c->Q = in->Q;
    c->max_citcom_cycle = in->max_citcom_cycle;
    c->max_cg_cycle = in->max_cg_cycle;
    c->cg_precondition = in->cg_precondition;
    c->check_v_only = in->check_v_only;
    if (!strcmp(in->solver, "jc")) c->solver = MCL_SOLVER_JC;
    else if (!strcmp(in->solver, "gs")) c->solver = MCL_SOLVER_GS;
    else if (!strcmp(in->solver, "cg")) c->solver = MCL_SOLVER_CG;
    else if (!strcmp(in->solver, "mg")) c->solver = MCL_SOLVER_MG;
    else c->solver = -1;
    c->auto_accuracy = in->auto_accuracy;
    c->zero_P = in->zero_P;
    c->AMR = in->AMR;
    c->mg_cycle = in->mg_cycle;
    c->mg_finest_iteration = in->mg_finest_iteration;
    c->mg_coarse_iteration = in->mg_coarse_iteration;
    c->schur = in->schur;
    c->precision = in->precision;
    c->warm_start = in->warm_start;
    c->inner_rtol = in->inner_rtol;
    c->exact_dots = in->exact_dots;
}

/* ---------------------------------------------------------------- quadtree */

/* Tree.c:391-456: Mcode = (loc << 4) | level; loc = Morton code with x on the
 * odd bits. */
static int mc_loc(int z) { return z >> 4; }
static int mc_lev(int z) { return z & 15; }
static int mc_make(int loc, int lev) { return (loc << 4) | lev; }

static void demorton(int z, int *x, int *y)
{
    int i;
    *x = *y = 0;
    for (i = 0; i < 32; i += 2) {
        *x |= (int)(((unsigned)z & (1u << (i + 1))) >> (i / 2 + 1));
        *y |= (int)(((unsigned)z & (1u << i)) >> (i / 2));
    }
}

static int log4i(int a)
{
    int b = 0;
    while (a >>= 2) b++;
    return b;
}

typedef struct {
    int *v;
    size_t n, cap;
} ivec;

static void iv_push(ivec *a, int x)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 1024;
        a->v = realloc(a->v, a->cap * sizeof(int));
    }
    a->v[a->n++] = x;
}

static int icmp(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

typedef struct {
    int L;          /* max_level */
    int max_nel;    /* 4^(L-1): elements of a rank's block at the finest level */
    ivec head, add;
} tree_t;

/* split_Mcode, Tree.c:327-338 */
static void split_code(const tree_t *t, int parent, int level, int child[4])
{
    int i, mul;
    level += 1;
    mul = t->max_nel >> (2 * level);
    for (i = 0; i < 4; i++) child[i] = parent + i * mul;
}

/* split_mul_lev, Tree.c:302-325: append the children at every depth. */
static void split_mul(tree_t *t, int node, int num, int level, int to_add)
{
    int child[4], i;
    if (num == 0) return;
    split_code(t, node, level, child);
    for (i = 0; i < 4; i++) iv_push(to_add ? &t->add : &t->head, child[i]);
    for (i = 0; i < 4; i++) split_mul(t, child[i], num - 1, level + 1, to_add);
}

/* refine_bottom / refine_top, Tree.c:459-558 */
static void refine_band(tree_t *t, int ip, int mode, int top)
{
    int start = ip * t->max_nel;
    size_t k, n = t->head.n;
    for (k = 0; k < n; k++) {
        int code = t->head.v[k], a, b;
        demorton(code - start, &a, &b);
        if (top) {
            if (b < 48) split_mul(t, code, 1, 4, 1);
            else if (b < 112) split_mul(t, code, 2, 4, 1);
            else if (b < 240) split_mul(t, code, 3, 4, 1);
            else split_mul(t, code, 4, 4, 1);
        } else if (mode == 2) {
            if (b <= 112) split_mul(t, code, 4, 5, 1);
            else if (b <= 240) split_mul(t, code, 3, 5, 1);
            else if (b <= 368) split_mul(t, code, 2, 5, 1);
            else split_mul(t, code, 1, 5, 1);
        } else if (mode == 3 || mode == 10) {
            if (b <= 48) split_mul(t, code, 4, 4, 1);
            else if (b <= 112) split_mul(t, code, 3, 4, 1);
            else if (b <= 240) split_mul(t, code, 2, 4, 1);
            else split_mul(t, code, 1, 4, 1);
        } else if (mode == 9) {
            if (b <= 48) split_mul(t, code, 4, 5, 1);
            else if (b <= 240) split_mul(t, code, 3, 5, 1);
            else if (b <= 368) split_mul(t, code, 2, 5, 1);
            else split_mul(t, code, 1, 5, 1);
        }
    }
    /* the refined list replaces the original one */
    free(t->head.v);
    t->head = t->add;
    t->add.v = NULL;
    t->add.n = t->add.cap = 0;
}

/* initial_refine, Tree.c:101-298, for virtual rank ip of an npz-tall layout. */
static int initial_refine(tree_t *t, int dev_mode, int ip, int npz)
{
    int M0 = ip * t->max_nel, L = t->L;
    int i, a, b, mul, width;
    switch (dev_mode) {
    case 0: split_mul(t, M0, L - 1, 0, 0); break;
    case 1:
        if (ip % 2 == 0) split_mul(t, M0, 1, 0, 0);
        else if (ip == 1) split_mul(t, M0, 2, 0, 0);
        else {
            split_mul(t, M0, 2, 0, 0);
            split_mul(t, 212, 1, 2, 0);
            split_mul(t, 220, 1, 2, 0);
            split_mul(t, 244, 1, 2, 0);
            split_mul(t, 252, 1, 2, 0);
        }
        break;
    case 2:
    case 9:
        split_mul(t, M0, 5, 0, 0);
        if (ip % 4 == 0) refine_band(t, ip, dev_mode, 0);
        break;
    case 3:
        split_mul(t, M0, 4, 0, 0);
        if (ip % 8 == 0) refine_band(t, ip, 3, 0);
        break;
    case 4: split_mul(t, M0, ip % 2 == 0 ? 2 : 1, 0, 0); break;
    case 5: split_mul(t, M0, ip % 2 == 0 ? 6 : 5, 0, 0); break;
    case 6:
        i = L - 1 - ip % npz;
        if (i < 4) i = 4;
        split_mul(t, M0, i, 0, 0);
        break;
    case 7: split_mul(t, M0, ip % npz == 0 ? L - 1 : L - 2, 0, 0); break;
    case 8: {
        static const int c2[] = {64, 80, 96, 112, 192, 208, 224, 240};
        static const int c3[] = {212, 220, 244, 252};
        split_mul(t, M0, 2, 0, 0);
        for (i = 0; i < 8; i++) split_mul(t, c2[i], 1, 2, 0);
        for (i = 0; i < 4; i++) split_mul(t, c3[i], 1, 3, 0);
        break;
    }
    case 10:
        split_mul(t, M0, 4, 0, 0);
        if (ip % 8 == 0) refine_band(t, ip, 10, 0);
        else if (ip % 8 == 7) refine_band(t, ip, 10, 1);
        break;
    case 11:
    case 12:
    case 14:
        split_mul(t, M0, L - 2, 0, 0);
        mul = 1 << (L - 1);
        width = dev_mode == 14 ? mul / 2 + 1 : mul / 8;
        for (i = 0; i < mul * mul; i += 4) {
            int hit;
            demorton(i, &a, &b);
            if (dev_mode == 11) hit = (b < width || b >= mul - width) || (a < width || a >= mul - width);
            else if (dev_mode == 12) hit = b >= mul - width;
            else hit = b <= width;
            if (hit) split_mul(t, i, 1, L - 2, 0);
        }
        break;
    case 13:
        split_mul(t, M0, L - 6, 0, 0);
        mul = 1 << (L - 1);
        width = mul / 64;
        for (i = 0; i < mul * mul; i += 1024) {
            demorton(i, &a, &b);
            if (b <= width) split_mul(t, i, 5, L - 6, 0);
            else if (b <= 4 * width) split_mul(t, i, 4, L - 6, 0);
            else if (b <= 8 * width) split_mul(t, i, 3, L - 6, 0);
            else if (b <= 12 * width) split_mul(t, i, 2, L - 6, 0);
            else split_mul(t, i, 1, L - 6, 0);
        }
        break;
    default: return 1;
    }
    return 0;
}

/* Sorted unique leaf codes of one virtual rank with their levels
 * (Tree.c:276-298 list_to_array, sort, remove_dup; append_level_to_Mcode
 * :340-358 derives each element's level from the gap to the next code, and the
 * last element's level from the gap to the previous one). */
static int rank_codes(int L, int dev_mode, int ip, int npz, ivec *out)
{
    tree_t t;
    size_t n, j, e;
    int max_nel_lev = 1 << (2 * (L - 1));
    memset(&t, 0, sizeof t);
    t.L = L;
    t.max_nel = max_nel_lev;
    if (initial_refine(&t, dev_mode, ip, npz)) {
        free(t.head.v);
        free(t.add.v);
        return 1;
    }
    qsort(t.head.v, t.head.n, sizeof(int), icmp);
    n = t.head.n;
    j = 0;
    for (e = 0; e < n; e++)
        if (e == n - 1 || t.head.v[e] != t.head.v[e + 1]) t.head.v[j++] = t.head.v[e];
    n = j;
    if (n == 0) {
        free(t.head.v);
        free(t.add.v);
        return 1;
    }
    for (e = 0; e + 1 < n; e++) {
        int lvl = log4i(max_nel_lev / (t.head.v[e + 1] - t.head.v[e]));
        t.head.v[e] = mc_make(t.head.v[e], lvl);
    }
    if (n >= 2) {
        int lvl = log4i(max_nel_lev / (t.head.v[n - 1] - mc_loc(t.head.v[n - 2])));
        t.head.v[n - 1] = mc_make(t.head.v[n - 1], lvl);
    } else {
        t.head.v[0] = mc_make(t.head.v[0], 0);
    }
    for (e = 0; e < n; e++) iv_push(out, t.head.v[e]);
    free(t.head.v);
    free(t.add.v);
    return 0;
}

/* project_Mcode, Tree.c:361-389 */
static void project_codes(const int *up, int nup, int ulev, ivec *out)
{
    int e = 0;
    while (e < nup) {
        int lvl = mc_lev(up[e]), loc = mc_loc(up[e]);
        if (lvl < ulev) e++;
        else {
            e += 4;
            lvl -= 1;
        }
        iv_push(out, mc_make(loc, lvl));
    }
}

static int *dup_int(const int *a, size_t n)
{
    int *p = malloc((n ? n : 1) * sizeof(int));
    if (n) memcpy(p, a, n * sizeof(int));
    return p;
}

static int bsearch_node(const int *node, int nno, int g)
{
    int lo = 0, hi = nno - 1;
    if (g == -1) return -1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (node[mid] == g) return mid;
        if (node[mid] < g) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

/* One level's connectivity, boundary conditions and equations. */
static int build_level(const mclh_input *in, const mclh_mesh *m, mclh_level *l, const int *codes, int nel,
                       char *err, int errlen)
{
    int L = m->L, gnoz = m->gnoz, gnox = m->gnox, e, a, i, d;
    int max_nel_top = 1 << (2 * (L - 1));
    int elx = 1 << (L - 1);
    ivec nodes = {0};
    l->L = L;
    l->nel = nel;
    l->gnox = gnox;
    l->gnoz = gnoz;
    l->Mcode = dup_int(codes, nel);
    l->ien = malloc(4 * (size_t)nel * sizeof(int));
    /* construct_ien, Construct_ien_id_lm.c:3-41 */
    for (e = 0; e < nel; e++) {
        int lvl = mc_lev(codes[e]), loc = mc_loc(codes[e]);
        int ip = loc / max_nel_top, ipx, ipz, x, z, node, step;
        ipx = ip / in->npz;
        ipz = ip - ipx * in->npz;
        demorton(loc - ip * max_nel_top, &x, &z);
        x += elx * ipx;
        z += elx * ipz;
        node = z + x * gnoz;
        step = 1 << (L - 1 - lvl);
        if (lvl > L - 1) return herr(err, errlen, "element %d deeper than max level", e);
        l->ien[0 * nel + e] = node;
        l->ien[1 * nel + e] = node + step;
        l->ien[2 * nel + e] = node + step + step * gnoz;
        l->ien[3 * nel + e] = node + step * gnoz;
    }
    /* node_index_from_ien, :43-67 */
    for (a = 0; a < 4; a++)
        for (e = 0; e < nel; e++) iv_push(&nodes, l->ien[a * nel + e]);
    qsort(nodes.v, nodes.n, sizeof(int), icmp);
    {
        size_t j = 0, k;
        for (k = 0; k < nodes.n; k++)
            if (k == nodes.n - 1 || nodes.v[k] != nodes.v[k + 1]) nodes.v[j++] = nodes.v[k];
        nodes.n = j;
    }
    l->nno = (int)nodes.n;
    l->node = nodes.v;
    {
        const int nno = l->nno;
        int P;
        /* node_position, Mesh.c:430-452 */
        l->X = malloc(2 * (size_t)nno * sizeof(double));
        for (i = 0; i < nno; i++) {
            int ix = l->node[i] / gnoz, iz = l->node[i] - ix * gnoz;
            l->X[i] = ix * m->dx[L - 1];
            l->X[nno + i] = iz * m->dz[L - 1];
        }
        /* temperature_bc, velocity_bc (Boundary_condition.c:10-110) */
        l->T_bc = malloc(nno);
        l->T_bc_val = calloc(nno, sizeof(double));
        l->v_bc = malloc(2 * (size_t)nno);
        l->v_bc_val = calloc(2 * (size_t)nno, sizeof(double));
        for (i = 0; i < nno; i++) {
            int g = l->node[i];
            l->T_bc[i] = 'N';
            if ((g + 1) % gnoz == 0) {
                l->T_bc[i] = in->top_T_bc;
                l->T_bc_val[i] = in->top_T_bc_val;
            } else if (g % gnoz == 0) {
                l->T_bc[i] = in->bot_T_bc;
                l->T_bc_val[i] = in->bot_T_bc_val;
            } else if (g < gnoz) {
                if (l->T_bc[i] != 'T') l->T_bc[i] = in->lef_T_bc;
                l->T_bc_val[i] = in->lef_T_bc_val;
            } else if (g > (gnox - 1) * gnoz - 1) {
                if (l->T_bc[i] != 'T') l->T_bc[i] = in->rig_T_bc;
                l->T_bc_val[i] = in->rig_T_bc_val;
            }
        }
        for (d = 0; d < 2; d++)
            for (i = 0; i < nno; i++) {
                int g = l->node[i];
                signed char *f = &l->v_bc[d * nno + i];
                double *fv = &l->v_bc_val[d * nno + i];
                *f = 'N';
                if ((g + 1) % gnoz == 0) {
                    *f = in->top_v_bc[d];
                    *fv = in->top_v_bc_val[d];
                } else if (g % gnoz == 0) {
                    *f = in->bot_v_bc[d];
                    *fv = in->bot_v_bc_val[d];
                }
                if (g < gnoz) {
                    if (*f != 'V') *f = in->lef_v_bc[d];
                    *fv = in->lef_v_bc_val[d];
                } else if (g > (gnox - 1) * gnoz - 1) {
                    if (*f != 'V') *f = in->rig_v_bc[d];
                    *fv = in->rig_v_bc_val[d];
                }
            }
        /* construct_v_id / construct_T_id, Construct_ien_id_lm.c:79-144 */
        l->id = malloc(2 * (size_t)nno * sizeof(int));
        l->veq_node = malloc(2 * (size_t)nno * sizeof(int));
        l->veq_dof = malloc(2 * (size_t)nno * sizeof(int));
        P = 0;
        for (i = 0; i < nno; i++)
            for (d = 0; d < 2; d++) {
                if (l->v_bc[d * nno + i] == 'V') l->id[d * nno + i] = -1;
                else {
                    l->id[d * nno + i] = P;
                    l->veq_node[P] = l->node[i];
                    l->veq_dof[P] = d;
                    P++;
                }
            }
        l->neq = P;
        l->heat_id = malloc((size_t)nno * sizeof(int));
        l->heat_eq_node = malloc((size_t)nno * sizeof(int));
        P = 0;
        for (i = 0; i < nno; i++) {
            if (l->T_bc[i] == 'T') l->heat_id[i] = -1;
            else {
                l->heat_id[i] = P;
                l->heat_eq_node[P] = l->node[i];
                P++;
            }
        }
        l->neqT = P;
        /* construct_v_lm / construct_T_lm, :146-196; element_area :198-218;
         * what_elements_node_in :220-244 */
        l->lm = malloc(8 * (size_t)nel * sizeof(int));
        l->heat_lm = malloc(4 * (size_t)nel * sizeof(int));
        l->area = malloc((size_t)nel * sizeof(double));
        l->p2n_nel = calloc(nno, sizeof(int));
        l->p2n_el = malloc(4 * (size_t)nno * sizeof(int));
        for (i = 0; i < 4 * nno; i++) l->p2n_el[i] = -1;
        for (e = 0; e < nel; e++) {
            int li[4];
            for (a = 0; a < 4; a++) {
                li[a] = bsearch_node(l->node, nno, l->ien[a * nel + e]);
                for (d = 0; d < 2; d++) l->lm[(d * 4 + a) * nel + e] = l->id[d * nno + li[a]];
                l->heat_lm[a * nel + e] = l->heat_id[li[a]];
                if (l->p2n_nel[li[a]] >= 4) return herr(err, errlen, "node in more than 4 elements");
                l->p2n_el[l->p2n_nel[li[a]] * nno + li[a]] = e;
                l->p2n_nel[li[a]]++;
            }
            l->area[e] = (l->X[li[2]] - l->X[li[0]]) * (l->X[nno + li[2]] - l->X[nno + li[0]]);
        }
        /* find_hangling_nodes, Connection.c:59-289 (single-rank part): the edge
         * midpoints of elements coarser than the level that exist as nodes. */
        {
            ivec hn = {0}, hc0 = {0}, hc1 = {0};
            for (e = 0; e < nel; e++) {
                int nd[4];
                if (mc_lev(codes[e]) == l->lev) continue;
                nd[0] = (l->ien[0 * nel + e] + l->ien[1 * nel + e]) / 2;
                nd[1] = (l->ien[1 * nel + e] + l->ien[2 * nel + e]) / 2;
                nd[2] = (l->ien[2 * nel + e] + l->ien[3 * nel + e]) / 2;
                nd[3] = (l->ien[3 * nel + e] + l->ien[0 * nel + e]) / 2;
                for (i = 0; i < 4; i++) {
                    if (bsearch_node(l->node, nno, nd[i]) >= 0) {
                        iv_push(&hn, nd[i]);
                        iv_push(&hc0, l->ien[((0 + i) % 4) * nel + e]);
                        iv_push(&hc1, l->ien[((1 + i) % 4) * nel + e]);
                    }
                }
            }
            l->nh = (int)hn.n;
            l->h_node = dup_int(hn.v, hn.n);
            l->h_ni = malloc((hn.n ? hn.n : 1) * sizeof(int));
            l->h_cnode = malloc(2 * (hn.n ? hn.n : 1) * sizeof(int));
            l->h_ci = malloc(2 * (hn.n ? hn.n : 1) * sizeof(int));
            for (i = 0; i < l->nh; i++) {
                l->h_ni[i] = bsearch_node(l->node, nno, hn.v[i]);
                l->h_cnode[i] = hc0.v[i];
                l->h_cnode[l->nh + i] = hc1.v[i];
                l->h_ci[i] = bsearch_node(l->node, nno, hc0.v[i]);
                l->h_ci[l->nh + i] = bsearch_node(l->node, nno, hc1.v[i]);
            }
            free(hn.v);
            free(hc0.v);
            free(hc1.v);
        }
    }
    return 0;
}

static void free_level(mclh_level *l)
{
    free(l->Mcode); free(l->ien); free(l->node); free(l->id); free(l->lm); free(l->heat_id); free(l->heat_lm);
    free(l->veq_node); free(l->veq_dof); free(l->heat_eq_node); free(l->v_bc); free(l->v_bc_val);
    free(l->T_bc); free(l->T_bc_val); free(l->X); free(l->area); free(l->p2n_nel); free(l->p2n_el);
    free(l->h_node); free(l->h_ni); free(l->h_cnode); free(l->h_ci);
    free(l->interp_n); free(l->interp_eq); free(l->proj_l); free(l->proj_u);
    memset(l, 0, sizeof *l);
}

void mclh_free_mesh(mclh_mesh *m)
{
    int k;
    if (!m) return;
    for (k = 0; k < m->nlev; k++) free_level(&m->lev[k]);
    free(m->lev);
    free(m);
}

/* Multigrid transfer maps between level k-1 (coarse) and k (fine), the
 * single-rank parts of build_interp_conn (Connection.c:708-1050) and
 * build_project_conn / setup_project_comm_frame (:1051-1350). */
static void build_mg_maps(mclh_level *lo, mclh_level *up);

/* All levels from the finest level's sorted leaf codes (takes ownership of
 * top.v): project down level by level (build_quad_tree, Tree.c:67-93), build
 * each level's structures and the multigrid transfer maps. */
static int mesh_from_codes(const mclh_input *in, ivec top, mclh_mesh **out, char *err, int errlen)
{
    mclh_mesh *m;
    int L = in->max_level, k;
    *out = NULL;
    m = calloc(1, sizeof *m);
    m->L = L;
    m->min_level = in->min_level;
    m->nlev = L - in->min_level;
    /* get_time_independent_variables, Mesh.c:396-427 */
    for (k = 0; k < L; k++) {
        int mul = 1 << k;
        m->dx[k] = (in->length / in->npx) / mul;
        m->dz[k] = (in->height / in->npz) / mul;
    }
    m->gnox = (1 << (L - 1)) * in->npx + 1;
    m->gnoz = (1 << (L - 1)) * in->npz + 1;
    m->lev = calloc(m->nlev, sizeof(mclh_level));
    {
        int **codes = calloc(m->nlev, sizeof(int *));
        int *ncodes = calloc(m->nlev, sizeof(int));
        codes[m->nlev - 1] = top.v;
        ncodes[m->nlev - 1] = (int)top.n;
        for (k = m->nlev - 1; k > 0; k--) {
            ivec lo = {0};
            project_codes(codes[k], ncodes[k], in->min_level + k, &lo);
            codes[k - 1] = lo.v;
            ncodes[k - 1] = (int)lo.n;
        }
        for (k = 0; k < m->nlev; k++) {
            m->lev[k].lev = in->min_level + k;
            if (build_level(in, m, &m->lev[k], codes[k], ncodes[k], err, errlen)) {
                for (int j = 0; j < m->nlev; j++) free(codes[j]);
                free(codes);
                free(ncodes);
                mclh_free_mesh(m);
                return 1;
            }
        }
        for (k = 0; k < m->nlev; k++) free(codes[k]);
        free(codes);
        free(ncodes);
    }
    /* transfer maps for every multi-level mesh (the reference builds them for
     * solver = "mg"; the mixed-precision path uses them as its multigrid) */
    for (k = 1; k < m->nlev; k++) build_mg_maps(&m->lev[k - 1], &m->lev[k]);
    *out = m;
    return 0;
}

static int check_dims(const mclh_input *in, char *err, int errlen)
{
    int L = in->max_level, np = in->npx * in->npz;
    if (L < 2 || L > 16) return herr(err, errlen, "level must be in [2,16]");
    if (in->min_level < 0 || in->min_level >= L) return herr(err, errlen, "bad mg_min_level");
    if (np < 1 || in->npx < 1 || in->npz < 1) return herr(err, errlen, "bad nprox/nproz");
    return 0;
}

int mclh_build_mesh(const mclh_input *in, mclh_mesh **out, char *err, int errlen)
{
    ivec top = {0};
    int L = in->max_level, ip, np = in->npx * in->npz;
    *out = NULL;
    if (check_dims(in, err, errlen)) return 1;
    for (ip = 0; ip < np; ip++)
        if (rank_codes(L, in->dev_mode, ip, in->npz, &top)) {
            free(top.v);
            return herr(err, errlen, "develop_mode %d not available", in->dev_mode);
        }
    qsort(top.v, top.n, sizeof(int), icmp);
    return mesh_from_codes(in, top, out, err, errlen);
}

/* ----------------------------------------------------------- adaptivity
 * The quadtree as a depth grid: the depth of the leaf covering each finest
 * cell, cells [ix * gz + iz] over the whole box (gx = 2^(L-1) npx columns,
 * gz = 2^(L-1) npz rows; rank ip = ipx npz + ipz owns columns ipx 2^(L-1)..
 * and rows ipz 2^(L-1)..).  Leaves are read back in each rank's Morton order,
 * so the codes come out sorted as the reference's Mcode lists. */
static int depth_to_codes(const mclh_input *in, int L, const unsigned char *depth, ivec *out)
{
    const int elx = 1 << (L - 1), max_nel = 1 << (2 * (L - 1)), np = in->npx * in->npz;
    const size_t gz = (size_t)elx * in->npz;
    for (int ip = 0; ip < np; ip++) {
        int ipx = ip / in->npz, ipz = ip - ipx * in->npz;
        int m = 0;
        while (m < max_nel) {
            int x, z, d, cells;
            demorton(m, &x, &z);
            d = depth[(size_t)(x + elx * ipx) * gz + (size_t)(z + elx * ipz)];
            if (d > L - 1) return 1;
            cells = 1 << (2 * (L - 1 - d));
            if (m % cells != 0) return 1;
            iv_push(out, mc_make(ip * max_nel + m, d));
            m += cells;
        }
    }
    return 0;
}

static void fill_depth(const mclh_input *in, int L, const int *codes, int nel, unsigned char *depth)
{
    const int elx = 1 << (L - 1), max_nel = 1 << (2 * (L - 1));
    const size_t gz = (size_t)elx * in->npz;
    for (int e = 0; e < nel; e++) {
        int lvl = mc_lev(codes[e]), loc = mc_loc(codes[e]);
        int ip = loc / max_nel, ipx = ip / in->npz, ipz = ip - ipx * in->npz, x, z;
        int s = 1 << (L - 1 - lvl);
        demorton(loc - ip * max_nel, &x, &z);
        x += elx * ipx;
        z += elx * ipz;
        for (int i = 0; i < s; i++)
            for (int j = 0; j < s; j++) depth[(size_t)(x + i) * gz + (size_t)(z + j)] = (unsigned char)lvl;
    }
}

int mclh_build_mesh_depth(const mclh_input *in, int depth, mclh_mesh **out, char *err, int errlen)
{
    int L = in->max_level;
    size_t ncell;
    unsigned char *grid;
    ivec top = {0};
    *out = NULL;
    if (check_dims(in, err, errlen)) return 1;
    if (depth < 0 || depth > L - 1) return herr(err, errlen, "depth %d outside [0, %d]", depth, L - 1);
    ncell = (size_t)(1 << (L - 1)) * in->npx * (size_t)(1 << (L - 1)) * in->npz;
    grid = malloc(ncell);
    memset(grid, depth, ncell);
    if (depth_to_codes(in, L, grid, &top)) {
        free(grid);
        free(top.v);
        return herr(err, errlen, "invalid depth grid");
    }
    free(grid);
    return mesh_from_codes(in, top, out, err, errlen);
}

/* Raise the leaf block covering cell (x, z) at depth d to depth d + 1. */
static void refine_block(unsigned char *depth, size_t gz, int L, int x, int z, int d)
{
    int s = 1 << (L - 1 - d), x0 = (x / s) * s, z0 = (z / s) * s;
    for (int i = 0; i < s; i++)
        for (int j = 0; j < s; j++) depth[(size_t)(x0 + i) * gz + (size_t)(z0 + j)] = (unsigned char)(d + 1);
}

int mclh_adapt_mesh(const mclh_input *in, const mclh_mesh *old, const double *T_old, int min_depth, int max_depth,
                    double refine_thr, double coarsen_thr, mclh_mesh **out, double **T_new, int *nrefined,
                    int *ncoarsened, char *err, int errlen)
{
    const mclh_level *f = &old->lev[old->nlev - 1];
    const int L = old->L, elx = 1 << (L - 1), nel = f->nel, gnoz = f->gnoz;
    const size_t gx = (size_t)elx * in->npx, gz = (size_t)elx * in->npz;
    const double hx = old->dx[L - 1], hz = old->dz[L - 1];
    unsigned char *od, *nd;
    double *crit;
    int nref = 0, ncoa = 0, changed;
    ivec top = {0};
    *out = NULL;
    *T_new = NULL;
    if (max_depth > L - 1 || min_depth < 0 || min_depth > max_depth)
        return herr(err, errlen, "depth range [%d, %d] outside [0, %d]", min_depth, max_depth, L - 1);
    od = malloc(gx * gz);
    nd = malloc(gx * gz);
    crit = malloc((size_t)(nel ? nel : 1) * sizeof(double));
    fill_depth(in, L, f->Mcode, nel, od);
    memcpy(nd, od, gx * gz);
    /* criterion |grad T| h at each leaf's centre (bilinear T from its corners) */
    for (int e = 0; e < nel; e++) {
        int lvl = mc_lev(f->Mcode[e]), s = 1 << (L - 1 - lvl);
        double t[4], ex = s * hx, ez = s * hz, gxv, gzv;
        for (int a = 0; a < 4; a++) t[a] = T_old[bsearch_node(f->node, f->nno, f->ien[a * nel + e])];
        gxv = ((t[2] + t[3]) - (t[0] + t[1])) / (2.0 * ex);
        gzv = ((t[1] + t[2]) - (t[0] + t[3])) / (2.0 * ez);
        crit[e] = sqrt(gxv * gxv + gzv * gzv) * (ex > ez ? ex : ez);
    }
    /* refinement */
    for (int e = 0; e < nel; e++) {
        int lvl = mc_lev(f->Mcode[e]);
        if (crit[e] > refine_thr && lvl < max_depth) {
            int n0 = f->ien[e], x = n0 / gnoz, z = n0 - x * gnoz;
            refine_block(nd, gz, L, x, z, lvl);
            nref++;
        }
    }
    /* coarsening of complete sibling families that all ask for it */
    for (int e = 0; e + 3 < nel; e++) {
        int lvl = mc_lev(f->Mcode[e]), loc = mc_loc(f->Mcode[e]);
        int cells = 1 << (2 * (L - 1 - lvl)), ok = 1;
        if (lvl <= min_depth || loc % (4 * cells) != 0) continue;
        for (int k = 0; k < 4 && ok; k++) {
            int c = f->Mcode[e + k];
            if (mc_lev(c) != lvl || mc_loc(c) != loc + k * cells || !(crit[e + k] < coarsen_thr)) ok = 0;
        }
        if (!ok) continue;
        {
            int n0 = f->ien[e], x = n0 / gnoz, z = n0 - x * gnoz, s = 2 << (L - 1 - lvl);
            int x0 = (x / s) * s, z0 = (z / s) * s, unchanged = 1;
            for (int i = 0; i < s && unchanged; i++)
                for (int j = 0; j < s; j++)
                    if (nd[(size_t)(x0 + i) * gz + (size_t)(z0 + j)] != lvl) { unchanged = 0; break; }
            if (!unchanged) continue;
            for (int i = 0; i < s; i++)
                for (int j = 0; j < s; j++) nd[(size_t)(x0 + i) * gz + (size_t)(z0 + j)] = (unsigned char)(lvl - 1);
            ncoa++;
            e += 3;
        }
    }
    /* 2:1 balance: a leaf next to one two or more levels deeper is refined */
    do {
        changed = 0;
        for (size_t x = 0; x < gx; x++)
            for (size_t z = 0; z < gz; z++) {
                int d = nd[x * gz + z];
                int dn = 0;
                if (x > 0 && nd[(x - 1) * gz + z] > dn) dn = nd[(x - 1) * gz + z];
                if (x + 1 < gx && nd[(x + 1) * gz + z] > dn) dn = nd[(x + 1) * gz + z];
                if (z > 0 && nd[x * gz + z - 1] > dn) dn = nd[x * gz + z - 1];
                if (z + 1 < gz && nd[x * gz + z + 1] > dn) dn = nd[x * gz + z + 1];
                if (dn > d + 1) {
                    refine_block(nd, gz, L, (int)x, (int)z, d);
                    changed = 1;
                }
            }
    } while (changed);
    free(crit);
    if (depth_to_codes(in, L, nd, &top)) {
        free(od);
        free(nd);
        free(top.v);
        return herr(err, errlen, "adapted depth grid is not a quadtree");
    }
    free(nd);
    if (mesh_from_codes(in, top, out, err, errlen)) {
        free(od);
        return 1;
    }
    /* temperature transfer: nodes of the old mesh keep their value; new nodes
     * take the bilinear interpolant of the old leaf that contains them */
    {
        const mclh_level *g = &(*out)->lev[(*out)->nlev - 1];
        double *T = malloc((size_t)g->nno * sizeof(double));
        for (int i = 0; i < g->nno; i++) {
            int nid = g->node[i], k = bsearch_node(f->node, f->nno, nid);
            if (k >= 0) {
                T[i] = T_old[k];
            } else {
                int ix = nid / gnoz, iz = nid - ix * gnoz;
                int cx = ix < (int)gx ? ix : (int)gx - 1, cz = iz < (int)gz ? iz : (int)gz - 1;
                int d = od[(size_t)cx * gz + (size_t)cz], s = 1 << (L - 1 - d);
                int x0 = (cx / s) * s, z0 = (cz / s) * s;
                double xi = (double)(ix - x0) / s, eta = (double)(iz - z0) / s;
                double t00 = T_old[bsearch_node(f->node, f->nno, z0 + x0 * gnoz)];
                double t01 = T_old[bsearch_node(f->node, f->nno, (z0 + s) + x0 * gnoz)];
                double t11 = T_old[bsearch_node(f->node, f->nno, (z0 + s) + (x0 + s) * gnoz)];
                double t10 = T_old[bsearch_node(f->node, f->nno, z0 + (x0 + s) * gnoz)];
                T[i] = (1 - xi) * (1 - eta) * t00 + (1 - xi) * eta * t01 + xi * eta * t11 + xi * (1 - eta) * t10;
            }
        }
        *T_new = T;
    }
    free(od);
    if (nrefined) *nrefined = nref;
    if (ncoarsened) *ncoarsened = ncoa;
    return 0;
}

void mclh_free_array(void *p) { free(p); }

/* sizes of the structs shared with bindings, for layout checks */
long mclh_sizeof(int which)
{
    switch (which) {
    case 0: return (long)sizeof(mclh_input);
    case 1: return (long)sizeof(mcl_control);
    case 2: return (long)sizeof(mcl_mesh_view);
    case 3: return (long)sizeof(mcl_stokes_stats);
    case 4: return (long)sizeof(mcl_heat_stats);
    case 5: return (long)sizeof(mclh_level);
    case 6: return (long)sizeof(mclh_mesh);
    default: return -1;
    }
}

static void build_mg_maps(mclh_level *lo, mclh_level *up)
{
    int i, d, e, a;
    /* projection: coarse eq <- coincident fine eq, in fine-node order */
    {
        ivec pl = {0}, pu = {0};
        for (i = 0; i < up->nno; i++) {
            int j = bsearch_node(lo->node, lo->nno, up->node[i]);
            if (j < 0) continue;
            for (d = 0; d < 2; d++) {
                int eq_l = lo->id[d * lo->nno + j];
                if (eq_l >= 0) {
                    iv_push(&pl, eq_l);
                    iv_push(&pu, up->id[d * up->nno + i]);
                }
            }
        }
        up->nproj = (int)pl.n;
        up->proj_l = pl.v ? pl.v : malloc(sizeof(int));
        up->proj_u = pu.v ? pu.v : malloc(sizeof(int));
    }
    /* interpolation: each fine eq <- mean of the coarse eqs of the coarse
     * element corners it lies on (1 coincident, 2 on an edge, 4 at a centre) */
    up->interp_n = calloc(up->neq ? up->neq : 1, sizeof(int));
    up->interp_eq = malloc(4 * (size_t)(up->neq ? up->neq : 1) * sizeof(int));
    for (i = 0; i < 4 * up->neq; i++) up->interp_eq[i] = -1;
    for (e = 0; e < lo->nel; e++) {
        int c[4], k, g;
        for (a = 0; a < 4; a++) c[a] = lo->ien[a * lo->nel + e];
        /* corners, edge midpoints, centre of the coarse element */
        for (k = 0; k < 9; k++) {
            int fine, src[4], ns = 0;
            if (k < 4) {
                fine = c[k];
                src[0] = c[k];
                ns = 1;
            } else if (k < 8) {
                int a0 = k - 4, a1 = (k - 3) % 4;
                fine = (c[a0] + c[a1]) / 2;
                src[0] = c[a0];
                src[1] = c[a1];
                ns = 2;
            } else {
                fine = (c[0] + c[2]) / 2;
                for (g = 0; g < 4; g++) src[g] = c[g];
                ns = 4;
            }
            {
                int fi = bsearch_node(up->node, up->nno, fine);
                if (fi < 0) continue;
                for (d = 0; d < 2; d++) {
                    int u = up->id[d * up->nno + fi];
                    if (u < 0 || up->interp_n[u] > 0) continue;
                    up->interp_n[u] = ns;
                    for (g = 0; g < ns; g++) {
                        int li = bsearch_node(lo->node, lo->nno, src[g]);
                        up->interp_eq[g * up->neq + u] = li >= 0 ? lo->id[d * lo->nno + li] : -1;
                    }
                }
            }
        }
    }
}

void mclh_fill_view(const mclh_mesh *m, int k, mcl_mesh_view *v)
{
    const mclh_level *l = &m->lev[k];
    memset(v, 0, sizeof *v);
    v->L = m->L;
    v->lev = l->lev;
    v->nel = l->nel;
    v->nno = l->nno;
    v->neq = l->neq;
    v->neqT = l->neqT;
    v->gnox = m->gnox;
    v->gnoz = m->gnoz;
    v->Mcode = l->Mcode;
    v->ien = l->ien;
    v->node = l->node;
    v->id = l->id;
    v->lm = l->lm;
    v->heat_id = l->heat_id;
    v->heat_lm = l->heat_lm;
    v->v_bc = l->v_bc;
    v->v_bc_val = l->v_bc_val;
    v->T_bc = l->T_bc;
    v->T_bc_val = l->T_bc_val;
    v->X = l->X;
    v->dx = m->dx;
    v->dz = m->dz;
    v->nh = l->nh;
    v->h_ni = l->h_ni;
    v->h_ci = l->h_ci;
    v->interp_n = l->interp_n;
    v->interp_eq = l->interp_eq;
    v->nproj = l->nproj;
    v->proj_l = l->proj_l;
    v->proj_u = l->proj_u;
}

/* ------------------------------------------------------- initial temperature */

double mclh_seeded_unit(unsigned long long seed, long long node)
{
    unsigned long long z = seed + (unsigned long long)node + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* initial_temperature, Initial_condition.c:3-135, for mesh level k.
 * T_ini: 0 conductive, 1 periodic, 2 random, 3 periodic + random, -1 file,
 * 4 (this build) conductive + seeded perturbation
 *     T = (H - z) + perturb * (2u - 1) * sin(pi (H - z) / H). */
int mclh_initial_temperature(const mclh_input *in, const mclh_mesh *m, int k, double *T, char *err, int errlen)
{
    const mclh_level *l = &m->lev[k];
    const double H = in->height, Lx = in->length;
    int i;
    for (i = 0; i < l->nno; i++) {
        double x = l->X[i], z = l->X[l->nno + i], r;
        switch (in->T_ini) {
        case 0:
            T[i] = H - z + in->T_mantle_perturb * sin(M_PI * (H - z) / H) * cos(in->T_wave_number * M_PI * x / Lx);
            break;
        case 1:
            T[i] = in->T_mantle + in->T_mantle_perturb * sin(M_PI * (H - z) / H) *
                                      cos(in->T_wave_number * M_PI * x / Lx);
            break;
        case 2:
            r = (1.0 * rand()) / (1.0 * RAND_MAX);
            T[i] = in->T_mantle + in->T_mantle_perturb * r;
            break;
        case 3:
            r = (1.0 * rand()) / (1.0 * RAND_MAX);
            T[i] = in->T_mantle + in->T_mantle_perturb * sin(M_PI * (H - z) / H) *
                                      cos(in->T_wave_number * M_PI * x / Lx);
            T[i] += r * 0.1 * in->T_mantle_perturb;
            break;
        case 4:
            r = 2.0 * mclh_seeded_unit(in->T_seed, l->node[i]) - 1.0;
            T[i] = (H - z) + in->T_mantle_perturb * r * sin(M_PI * z / H);
            break;
        case -1: {
            /* read_T_from_file, :171-195: one %lf per line in sorted node order */
            char fn[1200], buf[1000];
            FILE *fp;
            int j;
            if (k != m->nlev - 1) {
                T[i] = 0.0;
                break;
            }
            snprintf(fn, sizeof fn, "%s/%07d/temperature.%d.%07d", in->old_T_file, in->old_T_timestep, 0,
                     in->old_T_timestep);
            fp = fopen(fn, "r");
            if (!fp) return herr(err, errlen, "Cannot find file: %s", fn);
            for (j = 0; j < l->nno; j++) {
                if (!fgets(buf, sizeof buf, fp) || sscanf(buf, "%lf", &T[j]) != 1) {
                    fclose(fp);
                    return herr(err, errlen, "%s: short file", fn);
                }
            }
            fclose(fp);
            i = l->nno;
            break;
        }
        default: return herr(err, errlen, "Initial temperature is not set");
        }
    }
    /* apply_T_bc_values, :122-135 */
    for (i = 0; i < l->nno; i++)
        if (l->T_bc[i] == 'T') T[i] = l->T_bc_val[i];
    return 0;
}

/* ------------------------------------------------------------------- output */

/* Output.c:4-156: data_dir/%07d/<field>.<ip>.%07d, one %e value (or pair) per
 * node in sorted-node order; stress divided by 1e6. */
int mclh_write_output(const mclh_input *in, const mclh_level *l, int timestep, const double *T,
                      const double *stress, const double *vis, const double *V, char *err, int errlen)
{
    char dir[600], fn[800];
    FILE *fp;
    int i;
    snprintf(dir, sizeof dir, "%s", in->data_dir);
    mkdir(dir, 0755);
    snprintf(dir, sizeof dir, "%s/%07d", in->data_dir, timestep);
    if (mkdir(dir, 0755) != 0 && errno != EEXIST) return herr(err, errlen, "cannot create %s", dir);
#define OPEN(name)                                                                  \
    snprintf(fn, sizeof fn, "%s/%s.%d.%07d", dir, name, 0, timestep);                \
    fp = fopen(fn, "w");                                                            \
    if (!fp) return herr(err, errlen, "cannot write %s", fn);
    OPEN("coord");
    for (i = 0; i < l->nno; i++) fprintf(fp, "%e %e\n", l->X[i], l->X[l->nno + i]);
    fclose(fp);
    OPEN("temperature");
    for (i = 0; i < l->nno; i++) fprintf(fp, "%e\n", T[i]);
    fclose(fp);
    OPEN("stress");
    for (i = 0; i < l->nno; i++) fprintf(fp, "%e\n", stress[i] / 1e6);
    fclose(fp);
    OPEN("viscosity");
    for (i = 0; i < l->nno; i++) fprintf(fp, "%e\n", vis[i]);
    fclose(fp);
    OPEN("velocity");
    for (i = 0; i < l->nno; i++) fprintf(fp, "%+e %+e\n", V[i], V[l->nno + i]);
    fclose(fp);
#undef OPEN
    return 0;
}

/* ------------------------------------------------------------- diagnostics */

int mclh_diagnostics(const mclh_input *in, const mclh_level *l, const double *T, const double *V,
                     double *vrms, double *nu_top)
{
    const double g = 1.0 / sqrt(3.0);
    const double gp[4][2] = {{-g, -g}, {-g, g}, {g, g}, {g, -g}};
    double sum_u2 = 0.0, area = 0.0, flux = 0.0, width = 0.0;
    int e, a, q;
    for (e = 0; e < l->nel; e++) {
        int li[4];
        double dx, dz;
        for (a = 0; a < 4; a++) {
            int lo = 0, hi = l->nno - 1, key = l->ien[a * l->nel + e];
            while (lo <= hi) {
                int mid = (lo + hi) >> 1;
                if (l->node[mid] == key) { lo = mid; break; }
                if (l->node[mid] < key) lo = mid + 1; else hi = mid - 1;
            }
            li[a] = lo;
        }
        dx = l->X[li[2]] - l->X[li[0]];
        dz = l->X[l->nno + li[2]] - l->X[l->nno + li[0]];
        for (q = 0; q < 4; q++) {
            double xi = gp[q][0], et = gp[q][1];
            double N[4] = {0.25 * (1 - xi) * (1 - et), 0.25 * (1 - xi) * (1 + et), 0.25 * (1 + xi) * (1 + et),
                           0.25 * (1 + xi) * (1 - et)};
            double ux = 0.0, uz = 0.0;
            for (a = 0; a < 4; a++) {
                ux += N[a] * V[li[a]];
                uz += N[a] * V[l->nno + li[a]];
            }
            sum_u2 += (ux * ux + uz * uz) * 0.25 * dx * dz;
        }
        area += dx * dz;
        /* element touching the top: mean dT/dz over its top edge */
        if ((l->ien[1 * l->nel + e] + 1) % l->gnoz == 0) {
            double dTdz = 0.5 * ((T[li[1]] - T[li[0]]) + (T[li[2]] - T[li[3]])) / dz;
            flux += -dTdz * dx;
            width += dx;
        }
    }
    *vrms = sqrt(sum_u2 / area);
    {
        double dT = in->bot_T_bc_val - in->top_T_bc_val;
        *nu_top = (width > 0 && dT != 0) ? (flux / width) * in->height / dT : 0.0;
    }
    return 0;
}
