/*
 * ahop_oracle.h -- CPU restatement of the AdaptaHOP hot path (TEST
 * INFRASTRUCTURE ONLY: the checker used by tests, smoke() and the benchmark's
 * CPU-baseline leg; the product never links it).
 *
 * Every routine follows reference fortran/compute_adaptahop.f90 line by line
 * in IEEE double arithmetic with no contraction, so on the same inputs it
 * yields the reference's kNN lists (including its traversal-order tie
 * handling), densities, hop targets, group ids and saddle CSR bit for bit.
 * Indices in this API are 0-based; group ids are 1-based with 0 = inactive,
 * exactly as in the reference's liste_parts after find_local_maxima.
 */
#ifndef AHOP_ORACLE_H
#define AHOP_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Octree build (create_tree_structure f90:2967-3029, create_KDtree :3033-3221).
 * pos: (n,3) row-major == the reference's (3,n) Fortran array. */
orc_ctx *orc_create(int64_t n, const double *pos, const double *mass, int nvoisins, int nhop,
                    double rho_threshold, double xlong, int nlevelmax, int nthreads);
void orc_destroy(orc_ctx *c);
int64_t orc_ncells(const orc_ctx *c);

/* kNN of particle i (omp_find_nearest_parts :2841 + walk_tree2 :2867):
 * d2[0..k-1] ascending, ids[0..k-1] 0-based. */
void orc_knn(const orc_ctx *c, int64_t i, double *d2, int32_t *ids);

/* Density of every particle (compute_mean_density_and_np :682, omp_compute_density :2816).
 * Optionally also stores the first nhop neighbour ids (n x nhop, 0-based). Returns nactive. */
int64_t orc_density(const orc_ctx *c, double *density, int32_t *nbr_nhop);

/* Density of the m particles idx[0..m-1] (0-based), for sampled checks on large inputs. */
void orc_density_at(const orc_ctx *c, int64_t m, const int64_t *idx, double *out);

/* Hop targets and group ids (find_local_maxima :748-881).  hop: 0-based target of
 * each active particle (-1 inactive); gid: 1-based group id, 0 inactive;
 * densityg: density of each group's root (capacity n).  Returns ngroups. */
int32_t orc_maxima(const orc_ctx *c, const double *density, int32_t *hop, int32_t *gid,
                   double *densityg);

/* Saddle CSR (dev_compute_saddle_list :2433-2813).  offset[0..ngroups] 0-based;
 * entries sorted by decreasing saddle density within each group as the
 * reference leaves them.  Returns the entry count; fills isad/rho only when
 * cap >= count. */
int64_t orc_saddle(const orc_ctx *c, const double *density, const int32_t *gid, int32_t ngroups,
                   int32_t *offset, int32_t *isad, double *rho, int64_t cap);

/* O(N) brute-force periodic kNN in the reference distance arithmetic, used to
 * pin orc_knn on small inputs.  Ties are broken by ascending index. */
void orc_knn_brute(int64_t n, const double *pos, double xlong, int64_t i, int k, double *d2,
                   int32_t *ids);

/* Numerical Recipes index sort as the reference ships it (indexx :220-303), 0-based out. */
void orc_indexx(int32_t n, const double *arr, int32_t *indx);

#ifdef __cplusplus
}
#endif
#endif
