/*
 * ahop_oracle.c -- CPU restatement of the AdaptaHOP hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Tests, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it as the checker; the product library never links
 * or calls it.
 *
 * Follows reference fortran/compute_adaptahop.f90 statement by statement
 * (cited per routine).  Built with -ffp-contract=off so no a*b+c is fused,
 * matching the reference build on x86-64.  Internally the octree keeps the
 * reference's 1-based cell and particle numbering so that its traversal order,
 * and therefore its treatment of equal distances, is reproduced exactly.
 */
#include "ahop_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NPARTPERCELL 100      /* f90:91 */
#define BIGNUM 1.0e30         /* f90:94 */
/* f90:58 declares pi = 3.141592654 as a default-real literal: single rounded. */
static const double PI_REF = (double)3.141592654f;

struct orc_ctx {
    int64_t n;
    const double *pos;   /* (n,3) row-major: pos[3*(ipar-1)+d] */
    const double *mass;
    int k, nhop, nlevelmax, nthreads;
    double rho_thr, xlong, ylong, zlong, xyzlong;
    /* octree, 1-based cells */
    int64_t ncell, capcell;
    double *pos_cell;    /* 3*(cap+1) */
    double *size_cell;
    int32_t *mass_cell, *sister, *firstchild;
    int32_t *idpart;     /* 1-based slots, values 1-based particle ids */
    int32_t *idpart_tmp, *icid_tmp;
};

static const double POS_REF[8][3] = {  /* pos_ref_kdtree f90:473-481 */
    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
    {-1, -1, 1},  {1, -1, 1},  {-1, 1, 1},  {1, 1, 1}};

#define P(c, ip, d) ((c)->pos[3 * ((int64_t)(ip)-1) + (d)])

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }

static void grow_cells(orc_ctx *c, int64_t need)
{
    if (need <= c->capcell) return;
    int64_t cap = c->capcell;
    while (cap < need) cap = cap + (cap / 2) + 1024;
    c->pos_cell = realloc(c->pos_cell, sizeof(double) * 3 * (cap + 1));
    c->size_cell = realloc(c->size_cell, sizeof(double) * (cap + 1));
    c->mass_cell = realloc(c->mass_cell, sizeof(int32_t) * (cap + 1));
    c->sister = realloc(c->sister, sizeof(int32_t) * (cap + 1));
    c->firstchild = realloc(c->firstchild, sizeof(int32_t) * (cap + 1));
    c->capcell = cap;
}

/* create_KDtree f90:3033-3221 */
static void create_kdtree(orc_ctx *c, int nlevel, const double pc[3], int32_t np_node,
                          int32_t np_pos, int32_t idmother, double size_my)
{
    if (np_node <= 0) return;
    int64_t me = ++c->ncell;
    grow_cells(c, me);
    c->pos_cell[3 * me + 0] = pc[0];
    c->pos_cell[3 * me + 1] = pc[1];
    c->pos_cell[3 * me + 2] = pc[2];
    c->mass_cell[me] = np_node;
    c->size_cell[me] = size_my;
    c->sister[me] = 0;
    c->firstchild[me] = 0;
    if (idmother > 0) {
        c->sister[me] = c->firstchild[idmother];
        c->firstchild[idmother] = (int32_t)me;
    }
    if (np_node <= NPARTPERCELL || nlevel == c->nlevelmax) {
        c->firstchild[me] = -(np_pos + 1);
        return;
    }
    int32_t inc[8] = {0}, nsub[8];
    for (int32_t ip = np_pos + 1; ip <= np_pos + np_node; ++ip) {
        int32_t ipid = c->idpart[ip];
        int icid = 0;
        if (P(c, ipid, 0) >= pc[0]) icid += 1;
        if (P(c, ipid, 1) >= pc[1]) icid += 2;
        if (P(c, ipid, 2) >= pc[2]) icid += 4;
        c->icid_tmp[ip] = icid;
        inc[icid]++;
    }
    nsub[0] = 0;
    for (int j = 1; j < 8; ++j) nsub[j] = nsub[j - 1] + inc[j - 1];
    memset(inc, 0, sizeof inc);
    for (int32_t ip = np_pos + 1; ip <= np_pos + np_node; ++ip) {
        int icid = c->icid_tmp[ip];
        inc[icid]++;
        c->idpart_tmp[inc[icid] + nsub[icid] + np_pos] = c->idpart[ip];
    }
    for (int32_t ip = np_pos + 1; ip <= np_pos + np_node; ++ip) c->idpart[ip] = c->idpart_tmp[ip];
    double size_child = size_my * 0.5;
    for (int j = 0; j < 8; ++j) {
        if (inc[j] <= 0) continue;
        double po[3];
        po[0] = pc[0] + size_child * POS_REF[j][0];
        po[1] = pc[1] + size_child * POS_REF[j][1];
        po[2] = pc[2] + size_child * POS_REF[j][2];
        create_kdtree(c, nlevel + 1, po, inc[j], np_pos + nsub[j], (int32_t)me, size_child);
    }
}

orc_ctx *orc_create(int64_t n, const double *pos, const double *mass, int nvoisins, int nhop,
                    double rho_threshold, double xlong, int nlevelmax, int nthreads)
{
    if (n <= NPARTPERCELL || nvoisins < 1 || nvoisins >= n || nhop > nvoisins) return NULL;
    orc_ctx *c = calloc(1, sizeof *c);
    c->n = n;
    c->pos = pos;
    c->mass = mass;
    c->k = nvoisins;
    c->nhop = nhop;
    c->nlevelmax = nlevelmax;
    c->nthreads = nthreads > 0 ? nthreads : 1;
    c->rho_thr = rho_threshold;
    c->xlong = c->ylong = c->zlong = xlong;
    c->xyzlong = xlong * xlong * xlong;   /* sync_from_init_adaptahop f90:630 */
    c->idpart = malloc(sizeof(int32_t) * (n + 1));
    c->idpart_tmp = malloc(sizeof(int32_t) * (n + 1));
    c->icid_tmp = malloc(sizeof(int32_t) * (n + 1));
    for (int64_t i = 1; i <= n; ++i) c->idpart[i] = (int32_t)i;
    grow_cells(c, n > 1024 ? n / 4 : 1024);
    double pc[3] = {0.0, 0.0, 0.0};
    double sizeroot = dmax(dmax(xlong, xlong), xlong);   /* f90:3012 */
    create_kdtree(c, 0, pc, (int32_t)n, 0, 0, sizeroot * 0.5);
    free(c->idpart_tmp);
    free(c->icid_tmp);
    c->idpart_tmp = c->icid_tmp = NULL;
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    free(c->pos_cell);
    free(c->size_cell);
    free(c->mass_cell);
    free(c->sister);
    free(c->firstchild);
    free(c->idpart);
    free(c);
}

int64_t orc_ncells(const orc_ctx *c) { return c->ncell; }

/* walk_tree2 f90:2867-2962; dist2[0..k] with dist2[0]=0, iparnei[1..k] 1-based */
static void walk_tree2(const orc_ctx *c, int32_t icellidin, const double ph[3], double *dist2,
                       int32_t iparid, int32_t *iparnei)
{
    const int k = c->k;
    int32_t icid[9];
    double discell2[9];
    int32_t ic_id = c->firstchild[icellidin];
    int inc = 1;
    double r2max = dist2[k];
    discell2[0] = 0.0;
    for (int j = 1; j <= 8; ++j) discell2[j] = 1.0e30;
    while (ic_id != 0) {
        double sizec = c->size_cell[ic_id];
        double dx = fabs(c->pos_cell[3 * ic_id + 0] - ph[0]);
        double dy = fabs(c->pos_cell[3 * ic_id + 1] - ph[1]);
        double dz = fabs(c->pos_cell[3 * ic_id + 2] - ph[2]);
        dx = dmax(0.0, dmin(dx, c->xlong - dx) - sizec);
        dy = dmax(0.0, dmin(dy, c->ylong - dy) - sizec);
        dz = dmax(0.0, dmin(dz, c->zlong - dz) - sizec);
        double distance2 = dx * dx + dy * dy + dz * dz;
        if (distance2 < r2max) {
            int idist = inc - 1;
            while (discell2[idist] > distance2) {
                discell2[idist + 1] = discell2[idist];
                icid[idist + 1] = icid[idist];
                idist--;
            }
            discell2[idist + 1] = distance2;
            icid[idist + 1] = ic_id;
            inc++;
        }
        ic_id = c->sister[ic_id];
    }
    for (int ic = 1; ic <= inc - 1; ++ic) {
        int32_t out = icid[ic];
        if (c->firstchild[out] < 0) {
            if (discell2[ic] < r2max) {
                int32_t npp = -c->firstchild[out] - 1;
                for (int32_t i = npp + 1; i <= npp + c->mass_cell[out]; ++i) {
                    int32_t ipc = c->idpart[i];
                    if (ipc == iparid) continue;
                    double dx = fabs(P(c, ipc, 0) - ph[0]);
                    dx = dmax(0.0, dmin(dx, c->xlong - dx));
                    double d = dx * dx;
                    if (d >= r2max) continue;
                    double dy = fabs(P(c, ipc, 1) - ph[1]);
                    dy = dmax(0.0, dmin(dy, c->ylong - dy));
                    d = d + dy * dy;
                    if (d >= r2max) continue;
                    double dz = fabs(P(c, ipc, 2) - ph[2]);
                    dz = dmax(0.0, dmin(dz, c->zlong - dz));
                    d = d + dz * dz;
                    if (d < r2max) {
                        int idist = k - 1;
                        while (dist2[idist] > d) {
                            dist2[idist + 1] = dist2[idist];
                            iparnei[idist + 1] = iparnei[idist];
                            idist--;
                        }
                        dist2[idist + 1] = d;
                        iparnei[idist + 1] = ipc;
                        r2max = dist2[k];
                    }
                }
            }
        } else if (discell2[ic] < r2max) {
            walk_tree2(c, out, ph, dist2, iparid, iparnei);
            r2max = dist2[k];
        }
    }
}

/* omp_find_nearest_parts f90:2841-2862 (1-based ipar) */
static void find_nearest(const orc_ctx *c, int32_t ipar, double *dist2, int32_t *iparnei)
{
    double ph[3] = {P(c, ipar, 0), P(c, ipar, 1), P(c, ipar, 2)};
    dist2[0] = 0.0;
    for (int i = 1; i <= c->k; ++i) dist2[i] = BIGNUM;
    walk_tree2(c, 1, ph, dist2, ipar, iparnei);
}

void orc_knn(const orc_ctx *c, int64_t i, double *d2, int32_t *ids)
{
    double *dist2 = malloc(sizeof(double) * (c->k + 1));
    int32_t *nei = malloc(sizeof(int32_t) * (c->k + 1));
    find_nearest(c, (int32_t)(i + 1), dist2, nei);
    for (int j = 0; j < c->k; ++j) {
        d2[j] = dist2[j + 1];
        ids[j] = nei[j + 1] - 1;
    }
    free(dist2);
    free(nei);
}

/* spline f90:3270-3287 */
static double spline(double x)
{
    if (x <= 1.0) {
        double x2 = x * x;
        double x3 = x2 * x;
        return 1.0 - 1.5 * x2 + 0.75 * x3;
    } else if (x <= 2.0) {
        double x2 = (2.0 - x);
        double x3 = x2 * x2 * x2;
        return 0.25 * x3;
    }
    return 0.0;
}

/* omp_compute_density f90:2816-2838 */
static double sph_density(const orc_ctx *c, int32_t ipar, const double *dist2, const int32_t *iparnei)
{
    const int k = c->k;
    double r = sqrt(dist2[k]) * 0.5;
    double r3 = r * r * r;
    double unsr = 1.0 / r;
    double contrib = 0.0;
    for (int idist = 1; idist <= k - 1; ++idist)
        contrib = contrib + c->mass[iparnei[idist] - 1] * spline(sqrt(dist2[idist]) * unsr);
    return c->xyzlong * (contrib + c->mass[ipar - 1]) / (PI_REF * r3);
}

/* compute_mean_density_and_np f90:682-744 */
int64_t orc_density(const orc_ctx *c, double *density, int32_t *nbr_nhop)
{
    int64_t nactive = 0;
    const int k = c->k;
#pragma omp parallel num_threads(c->nthreads) reduction(+ : nactive)
    {
        double *dist2 = malloc(sizeof(double) * (k + 1));
        int32_t *nei = malloc(sizeof(int32_t) * (k + 1));
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < c->n; ++i) {
            int32_t ipar = (int32_t)(i + 1);
            find_nearest(c, ipar, dist2, nei);
            density[i] = sph_density(c, ipar, dist2, nei);
            if (density[i] > c->rho_thr) nactive++;
            if (nbr_nhop)
                for (int j = 0; j < c->nhop; ++j) nbr_nhop[i * c->nhop + j] = nei[j + 1] - 1;
        }
        free(dist2);
        free(nei);
    }
    return nactive;
}

/* density of selected particles only (same arithmetic as orc_density) */
void orc_density_at(const orc_ctx *c, int64_t m, const int64_t *idx, double *out)
{
    const int k = c->k;
#pragma omp parallel num_threads(c->nthreads)
    {
        double *dist2 = malloc(sizeof(double) * (k + 1));
        int32_t *nei = malloc(sizeof(int32_t) * (k + 1));
#pragma omp for schedule(dynamic, 16)
        for (int64_t q = 0; q < m; ++q) {
            int32_t ipar = (int32_t)(idx[q] + 1);
            find_nearest(c, ipar, dist2, nei);
            out[q] = sph_density(c, ipar, dist2, nei);
        }
        free(dist2);
        free(nei);
    }
}

/* find_local_maxima f90:748-881 */
int32_t orc_maxima(const orc_ctx *c, const double *density, int32_t *hop, int32_t *gid,
                   double *densityg)
{
    const int64_t n = c->n;
    const int k = c->k;
    int32_t *adapt = calloc(n + 1, sizeof(int32_t));    /* idpart_adapt, 1-based */
    int32_t *liste = calloc(n + 1, sizeof(int32_t));    /* liste_parts, 1-based */
#pragma omp parallel num_threads(c->nthreads)
    {
        double *dist2 = malloc(sizeof(double) * (k + 1));
        int32_t *nei = malloc(sizeof(int32_t) * (k + 1));
#pragma omp for schedule(guided, 256)
        for (int64_t i = 1; i <= n; ++i) {
            if (density[i - 1] <= c->rho_thr) continue;
            find_nearest(c, (int32_t)i, dist2, nei);
            double densest = density[i - 1];
            int32_t iparsel = (int32_t)i;
            for (int idist = 1; idist <= c->nhop; ++idist) {
                int32_t iparid = nei[idist];
                if (density[iparid - 1] > densest) {
                    iparsel = iparid;
                    densest = density[iparid - 1];
                } else if (density[iparid - 1] == densest) {
                    iparsel = iparsel < iparid ? iparsel : iparid;
                }
            }
            adapt[i] = iparsel;
        }
        free(dist2);
        free(nei);
    }
    if (hop)
        for (int64_t i = 1; i <= n; ++i) hop[i - 1] = density[i - 1] > c->rho_thr ? adapt[i] - 1 : -1;
    int32_t ngroups = 0;
    for (int64_t i = 1; i <= n; ++i) {
        if (density[i - 1] <= c->rho_thr) continue;
        if (adapt[i] == i) adapt[i] = -(++ngroups);
    }
    for (int64_t i = 1; i <= n; ++i) {
        if (density[i - 1] <= c->rho_thr) continue;
        int32_t iparid = adapt[i];
        if (iparid < 0) {
            int32_t groupid = -iparid;
            if (densityg) densityg[groupid - 1] = density[i - 1];
            liste[i] = groupid;
        } else {
            int32_t groupid = 0;
            while (iparid > 0) {
                if (liste[iparid] > 0) {
                    groupid = liste[iparid];
                    break;
                }
                iparid = adapt[iparid];
            }
            if (iparid < 0) groupid = -iparid;
            liste[i] = groupid;
        }
    }
    for (int64_t i = 1; i <= n; ++i) gid[i - 1] = liste[i];
    free(adapt);
    free(liste);
    return ngroups;
}

/* indexx f90:220-303 (1-based internally, as shipped: the insertion pass scans down to 1) */
void orc_indexx(int32_t n, const double *arr0, int32_t *indx0)
{
    enum { M = 7, NSTACK = 50 };
    const double *arr = arr0 - 1;
    int32_t *indx = malloc(sizeof(int32_t) * (n + 1));
    int32_t istack[NSTACK + 1];
    int32_t i, indxt, ir, itemp, j, jstack, kk, l;
    double a;
    for (j = 1; j <= n; ++j) indx[j] = j;
    jstack = 0;
    l = 1;
    ir = n;
    for (;;) {
        if (ir - l < M) {
            for (j = l + 1; j <= ir; ++j) {
                indxt = indx[j];
                a = arr[indxt];
                for (i = j - 1; i >= 1; --i) {
                    if (arr[indx[i]] <= a) break;
                    indx[i + 1] = indx[i];
                }
                if (i < 1) i = 0;
                indx[i + 1] = indxt;
            }
            if (jstack == 0) break;
            ir = istack[jstack];
            l = istack[jstack - 1];
            jstack -= 2;
        } else {
            kk = (l + ir) / 2;
            itemp = indx[kk];
            indx[kk] = indx[l + 1];
            indx[l + 1] = itemp;
            if (arr[indx[l + 1]] > arr[indx[ir]]) {
                itemp = indx[l + 1]; indx[l + 1] = indx[ir]; indx[ir] = itemp;
            }
            if (arr[indx[l]] > arr[indx[ir]]) {
                itemp = indx[l]; indx[l] = indx[ir]; indx[ir] = itemp;
            }
            if (arr[indx[l + 1]] > arr[indx[l]]) {
                itemp = indx[l + 1]; indx[l + 1] = indx[l]; indx[l] = itemp;
            }
            i = l + 1;
            j = ir;
            indxt = indx[l];
            a = arr[indxt];
            for (;;) {
                do { i++; } while (arr[indx[i]] < a);
                do { j--; } while (arr[indx[j]] > a);
                if (j < i) break;
                itemp = indx[i]; indx[i] = indx[j]; indx[j] = itemp;
            }
            indx[l] = indx[j];
            indx[j] = indxt;
            jstack += 2;
            if (ir - i + 1 >= j - l) {
                istack[jstack] = ir;
                istack[jstack - 1] = i;
                ir = j - 1;
            } else {
                istack[jstack] = j - 1;
                istack[jstack - 1] = l;
                l = i;
            }
        }
    }
    for (j = 1; j <= n; ++j) indx0[j - 1] = indx[j] - 1;
    free(indx);
}

/* dev_compute_saddle_list f90:2433-2813, run serially (each group's list depends
 * only on its own member walk, so the OpenMP schedule does not change it). */
int64_t orc_saddle(const orc_ctx *c, const double *density, const int32_t *gid, int32_t ngroups,
                   int32_t *offset, int32_t *isad, double *rho, int64_t cap)
{
    const int64_t n = c->n;
    const int k = c->k;
    /* member linked lists in index order, prepended (f90:853-861) */
    int32_t *firstpart = calloc(ngroups + 1, sizeof(int32_t));
    int32_t *next = calloc(n + 1, sizeof(int32_t));
    for (int64_t i = 1; i <= n; ++i) {
        if (density[i - 1] <= c->rho_thr) continue;
        int32_t g = gid[i - 1];
        if (g > 0) {
            next[i] = firstpart[g];
            firstpart[g] = (int32_t)i;
        }
    }
    int32_t *nhnei = calloc(ngroups + 1, sizeof(int32_t));
    int32_t *off = calloc(ngroups + 2, sizeof(int32_t));
    char *touch = calloc(ngroups + 1, 1);
    int32_t *listg = calloc(ngroups + 1, sizeof(int32_t));
    double *dist2 = malloc(sizeof(double) * (k + 1));
    int32_t *nei = malloc(sizeof(int32_t) * (k + 1));
    int32_t max_neig = 0;
    for (int32_t g1 = 1; g1 <= ngroups; ++g1) {          /* first pass :2495-2531 */
        int32_t ineig = 0;
        for (int32_t ip1 = firstpart[g1]; ip1 > 0; ip1 = next[ip1]) {
            find_nearest(c, ip1, dist2, nei);
            for (int idist = 1; idist <= c->nhop; ++idist) {
                int32_t g2 = gid[nei[idist] - 1];
                if (g2 > 0 && g2 != g1 && !touch[g2]) {
                    ineig++;
                    touch[g2] = 1;
                    listg[ineig] = g2;
                }
            }
        }
        for (int32_t q = 1; q <= ineig; ++q) touch[listg[q]] = 0;
        nhnei[g1] = ineig;
        if (ineig > max_neig) max_neig = ineig;
    }
    off[1] = 1;                                          /* CSR offsets :2537-2542 */
    for (int32_t g1 = 1; g1 <= ngroups; ++g1) off[g1 + 1] = off[g1] + nhnei[g1];
    int64_t total = (int64_t)off[ngroups + 1] - 1;
    int32_t *isad_all = malloc(sizeof(int32_t) * (total + 1));
    double *rho_all = malloc(sizeof(double) * (total + 1));
    double *rho_thr_buf = malloc(sizeof(double) * (max_neig + 1));
    for (int32_t g1 = 1; g1 <= ngroups; ++g1) {          /* second pass :2573-2632 */
        int32_t neig = nhnei[g1];
        if (neig <= 0) continue;
        int32_t ineig = 0, base = off[g1];
        for (int32_t ip1 = firstpart[g1]; ip1 > 0; ip1 = next[ip1]) {
            find_nearest(c, ip1, dist2, nei);
            double d1 = density[ip1 - 1];
            for (int idist = 1; idist <= c->nhop; ++idist) {
                int32_t ip2 = nei[idist];
                int32_t g2 = gid[ip2 - 1];
                if (g2 > 0 && g2 != g1) {
                    double d2v = density[ip2 - 1];
                    double s = d1 < d2v ? d1 : d2v;
                    if (!touch[g2]) {
                        ineig++;
                        touch[g2] = 1;
                        listg[g2] = ineig;
                        rho_thr_buf[ineig] = s;
                        isad_all[base + ineig - 1] = g2;
                    } else {
                        int32_t q = listg[g2];
                        rho_thr_buf[q] = rho_thr_buf[q] > s ? rho_thr_buf[q] : s;
                    }
                }
            }
        }
        for (int32_t q = 1; q <= ineig; ++q) rho_all[base + q - 1] = rho_thr_buf[q];
        for (int32_t q = 1; q <= ineig; ++q) touch[isad_all[base + q - 1]] = 0;
    }
    for (int32_t g1 = 1; g1 <= ngroups; ++g1) {          /* symmetry :2650-2688 */
        int32_t neig = nhnei[g1];
        int32_t base = off[g1];
        for (int32_t in1 = 1; in1 <= neig; ++in1) {
            int32_t idx1 = base + in1 - 1;
            int exist = 0;
            int32_t g2 = isad_all[idx1];
            if (g2 > 0) {
                int32_t base2 = off[g2];
                for (int32_t in2 = 1; in2 <= nhnei[g2]; ++in2) {
                    int32_t idx2 = base2 + in2 - 1;
                    if (isad_all[idx2] == g1) {
                        exist = 1;
                        double s = rho_all[idx2] < rho_all[idx1] ? rho_all[idx2] : rho_all[idx1];
                        rho_all[idx2] = s;
                        rho_all[idx1] = s;
                        break;
                    }
                }
            }
            if (!exist) isad_all[idx1] = 0;
        }
    }
    /* compact + sort decreasing :2702-2791 */
    int64_t newtotal = 0;
    int32_t *newcnt = calloc(ngroups + 1, sizeof(int32_t));
    for (int32_t g1 = 1; g1 <= ngroups; ++g1)
        for (int32_t q = 0; q < nhnei[g1]; ++q)
            if (isad_all[off[g1] + q] > 0) newcnt[g1]++;
    offset[0] = 0;
    for (int32_t g1 = 1; g1 <= ngroups; ++g1) offset[g1] = offset[g1 - 1] + newcnt[g1];
    newtotal = offset[ngroups];
    if (cap >= newtotal) {
        double *rs = malloc(sizeof(double) * (max_neig + 1));
        int32_t *is = malloc(sizeof(int32_t) * (max_neig + 1));
        int32_t *ix = malloc(sizeof(int32_t) * (max_neig + 1));
        for (int32_t g1 = 1; g1 <= ngroups; ++g1) {
            int32_t m = 0;
            for (int32_t q = 0; q < nhnei[g1]; ++q) {
                int32_t g2 = isad_all[off[g1] + q];
                if (g2 > 0) {
                    rs[m] = rho_all[off[g1] + q];
                    is[m] = g2;
                    m++;
                }
            }
            if (m > 0) {
                orc_indexx(m, rs, ix);
                int32_t ob = offset[g1 - 1];
                for (int32_t q = 0; q < m; ++q) {   /* store decreasing */
                    isad[ob + m - 1 - q] = is[ix[q]];
                    rho[ob + m - 1 - q] = rs[ix[q]];
                }
            }
        }
        free(rs);
        free(is);
        free(ix);
    }
    free(newcnt);
    free(rho_thr_buf);
    free(isad_all);
    free(rho_all);
    free(dist2);
    free(nei);
    free(listg);
    free(touch);
    free(off);
    free(nhnei);
    free(next);
    free(firstpart);
    return newtotal;
}

void orc_knn_brute(int64_t n, const double *pos, double xlong, int64_t i, int k, double *d2,
                   int32_t *ids)
{
    for (int j = 0; j < k; ++j) {
        d2[j] = BIGNUM;
        ids[j] = -1;
    }
    const double *pi = pos + 3 * i;
    for (int64_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const double *pj = pos + 3 * j;
        double d = 0.0;
        for (int a = 0; a < 3; ++a) {
            double dx = fabs(pj[a] - pi[a]);
            dx = dmax(0.0, dmin(dx, xlong - dx));
            d = d + dx * dx;
        }
        if (d < d2[k - 1]) {
            int q = k - 1;
            while (q > 0 && d2[q - 1] > d) {
                d2[q] = d2[q - 1];
                ids[q] = ids[q - 1];
                q--;
            }
            d2[q] = d;
            ids[q] = (int32_t)j;
        }
    }
}
