/*
 * amdahop.h -- C ABI of the MI355X-native AdaptaHOP hot path (libamdahop.so).
 *
 * The first group of entry points is the drop-in for the reference's f2py module
 * compute_adaptahop.neikdtree (fortran/compute_adaptahop.pyf:1-70): the same five
 * calls and five readable arrays, with the same argument meaning, called the same
 * way by src/compute_neiKDtree_mod.py.  The differences are the error behaviour --
 * every call returns 0 on success and a nonzero status (message in
 * ahop_last_error()) where the Fortran executes STOP -- and that array outputs are
 * read through accessor functions instead of module variables.
 *
 * Plain C types only: pointers, sizes, doubles.  Not re-entrant (one module state
 * per process, like the Fortran module), synchronous (outputs are host-visible when a
 * call returns).
 */
#ifndef AMDAHOP_H
#define AMDAHOP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- drop-in surface of compute_adaptahop.neikdtree --------------------------------- */

/* replaces neikdtree.sync_from_change_pos (compute_adaptahop.pyf:15-22, f90:576-589) */
int ahop_sync_from_change_pos(int64_t npart, int32_t nusedpart, double epsilon,
                              double fudgepsilon, double xlong, double boxsize2);

/* replaces neikdtree.sync_from_init_adaptahop (.pyf:24-47, f90:591-641) */
int ahop_sync_from_init_adaptahop(int64_t npart, int32_t nmembthresh, int32_t nmembers,
                                  double omegal, double omega_lambda_f, double omega0,
                                  double omega_f, double aexp_max, double af, double hubble,
                                  double h_f, double boxsize2, double lf, double xlong,
                                  double ylong, double zlong, double xlongs2, double ylongs2,
                                  double zlongs2, double hub_pt, double aexp,
                                  const double *pos_shift /* [3] */);

/* replaces neikdtree.sync_others (.pyf:49-64, f90:643-678); method is not NUL-terminated */
int ahop_sync_others(int32_t verbose, int32_t megaverbose, int32_t npart, int32_t nbpes,
                     double rho_threshold, double massp, double boxsize, int32_t nhop,
                     int32_t nvoisins, double fudge, double alphap, const char *method,
                     int32_t method_len, int32_t nlevelmax, int32_t optimize_nodes);

/* replaces neikdtree.compute_adaptahop(pos_in, mass_in) (.pyf:10-13, f90:495-543).
 * pos_in is the (3, npart) Fortran-ordered array (i.e. (npart,3) row-major), host memory. */
int ahop_compute_adaptahop(const double *pos_in, int64_t pos_rows, int64_t pos_cols,
                           const double *mass_in, int64_t mass_len);

/* replaces neikdtree.close (.pyf:66-67, f90:3227-3264) */
int ahop_close(void);

/* readable module arrays of neikdtree (.pyf:4-8); NULL / 0 when not allocated */
const int32_t *ahop_liste_parts(int64_t *len);      /* liste_parts(nusedpart) */
const double *ahop_density(int64_t *len);           /* density(npart) */
int32_t ahop_nnodes(void);                          /* nnodes */
const double *ahop_real_table(int64_t *ncols);      /* real_table(8, nnodes), column-major */
const int32_t *ahop_integer_table(int64_t *ncols);  /* integer_table(6, nnodes), column-major */

const char *ahop_last_error(void);

/* ---- extensions behind the same surface (no counterpart in the f2py module) ---------- */

/* 1: FP32 box-relative search with certified FP64 re-rank (default); 0: FP64 throughout. */
int ahop_set_precision(int32_t mixed);
/* particles per device-tree leaf bucket, 1..64 (default 64) */
int ahop_set_bucket(int32_t bucket);

/* Stage outputs of the last ahop_compute_adaptahop, original particle order:
 * group ids of find_local_maxima (f90:797-861) and 0-based hop targets (-1 inactive). */
int32_t ahop_ngroups(void);
int ahop_get_groups(int32_t *gid, int32_t *hop, int64_t n);
/* saddle CSR of dev_compute_saddle_list (f90:2433-2813): offset[ngroups+1] 0-based,
 * isad 1-based group ids, rho saddle densities, decreasing within each group. */
int64_t ahop_nsaddle(void);
int ahop_get_saddle(int32_t *offset, int32_t *isad, double *rho);

/* Hot-path parameters for the device-resident entry point. */
typedef struct {
    int32_t nvoisins, nhop;
    double rho_threshold;
    double xlong, ylong, zlong;
    int32_t mixed;       /* as ahop_set_precision */
    int32_t bucket;      /* as ahop_set_bucket */
    int32_t saddle;      /* also extract saddle edges */
} ahop_hot_params;

typedef struct {
    int64_t nactive;
    int32_t ngroups;
    int64_t nfallback;   /* mixed path: queries re-run in FP64 */
    int64_t nsaddle;
    float t_sort, t_tree, t_knn, t_fallback, t_hop, t_saddle, t_total;  /* ms, HIP events */
    /* work counters of the main kNN launch: tree nodes visited, leaf buckets scanned,
     * candidate distances (per lane), heap-insertion steps, heap insertions, waves */
    int64_t knn_nodes, knn_buckets, knn_candidates, knn_insert_steps, knn_inserts, knn_waves;
    /* mixed path, k <= 64: the largest per-query candidate log, the queries whose log
     * overflowed and those whose FP32 order of the k nearest differed from the FP64 order
     * (both counted in nfallback) */
    int64_t knn_log_max, knn_log_overflow, knn_misorder;
} ahop_run_info;

/* kNN + SPH density + hop + group ids (+ saddle edges) on device-resident inputs:
 * d_pos (n,3) row-major doubles and d_mass (n) in device memory.  Writes density and
 * 1-based group ids in original particle order to d_density / d_gid (device pointers,
 * either may be NULL).  Synchronous.  Replaces the OpenMP stages compute_mean_density_and_np
 * (f90:682), find_local_maxima (f90:748) and dev_compute_saddle_list (f90:2433). */
int ahop_run_device(const double *d_pos, const double *d_mass, int64_t n,
                    const ahop_hot_params *params, double *d_density, int32_t *d_gid,
                    ahop_run_info *info);
/* Host copies of the saddle CSR of the last ahop_run_device (sizes from its info). */
int ahop_device_saddle(int32_t *offset, int32_t *isad, double *rho);
/* Hop targets of the last ahop_run_device, copied to d_hop (n int32, device pointer): the
 * 0-based original index each particle hops to (find_local_maxima f90:774-795: the densest of
 * itself and its first nhop neighbours, ties to the smaller index), -1 where inactive.  n must
 * be that run's particle count.  Synchronous. */
int ahop_device_hop(int32_t *d_hop, int64_t n);

/* kNN + SPH density alone, keeping every particle's neighbour row: d_density (n) and
 * d_nbr (n x nvoisins int32, original 0-based ids in ascending distance, self excluded),
 * both device pointers in original order (params->nhop and ->saddle are ignored), and the
 * Morton order of the particles to d_order (n int32, sorted position -> original index; may
 * be NULL).  The
 * per-particle lists the reference builds in omp_find_nearest_parts / walk_tree2
 * (f90:2841-2962, iparnei) and sums in omp_compute_density (f90:2816-2838), exported so a
 * caller can freeze them (e.g. for ahop_density_adjoint_device). */
int ahop_knn_device(const double *d_pos, const double *d_mass, int64_t n,
                    const ahop_hot_params *params, double *d_density, int32_t *d_nbr,
                    int32_t *d_order, ahop_run_info *info);

/* Adjoint of the SPH density of omp_compute_density (f90:2816-2838) with the neighbour rows
 * frozen: given d_grad_rho (n) = dLoss/drho, writes dLoss/dpos to d_grad_pos (n,3) and
 * dLoss/dmass to d_grad_mass (n).  d_nbr (n x k) as ahop_knn_device writes it.  d_order (n,
 * may be NULL): a permutation of the particles, the d_order of ahop_knn_device; with it the
 * neighbours' terms are first summed per workgroup over a window of that order, which cuts
 * the global atomics several-fold.  No
 * counterpart in the reference (which has no gradients); it differentiates the density the
 * reference computes.  FP64 throughout; neighbour contributions are added with atomics,
 * so repeated calls agree to rounding.  *ms = device time (may be NULL).  Synchronous. */
int ahop_density_adjoint_device(const double *d_pos, const double *d_mass, int64_t n, int32_t k,
                                const int32_t *d_nbr, const int32_t *d_order,
                                const double *d_grad_rho, double xlong,
                                double ylong, double zlong, double *d_grad_pos,
                                double *d_grad_mass, float *ms);

/* Allocation-free variant of ahop_density_adjoint_device.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * or training loops, harnesses that forbid hipMalloc).  Same arguments, checks and outputs,
 * without the timing; it allocates nothing (no device memory, stream or event) and runs on
 * the device's null stream.  Its scratch -- an error counter and the inverse of d_order -- is a
 * fixed array in the library's device memory (about 16 MiB, held by each GPU the library runs
 * on until the process ends), so n is at most AHOP_ADJOINT_NOALLOC_MAX; larger n is an error.
 * Every call on a GPU shares that array: calls that overlap on one GPU (from other host
 * threads) race.  Synchronous. */
#define AHOP_ADJOINT_NOALLOC_MAX 4194304
int ahop_density_adjoint_device_noalloc(const double *d_pos, const double *d_mass, int64_t n,
                                        int32_t k, const int32_t *d_nbr, const int32_t *d_order,
                                        const double *d_grad_rho, double xlong, double ylong,
                                        double zlong, double *d_grad_pos, double *d_grad_mass);

/* ---- particles grouped by structure ------------------------------------------------- */

/* Members of every structure in one call, on device arrays.  d_node_of_part (n): the node id
 * of each particle (liste_parts, 1-based, 0 = in no node).  d_node_to_struct (nnodes, may be
 * NULL): 1-based structure id of each node, as select_with_MS_method_1319 builds it
 * (src/compute_neiKDtree_mod.py:341-348); with NULL the labels are the node ids and nstruct
 * must equal nnodes.  Outputs: d_struct_of_part (n, may be NULL) = the relabelled particles
 * (:376-379); d_pids (n) = 0-based particle indices sorted by label, ascending within a label;
 * d_offsets (nstruct + 2, int64): label l (0 = unassigned) owns d_pids[off[l] .. off[l+1]).
 * Replaces the relabelling of :376-379 and the grouping of compute_adaptahop_131 (:109-119:
 * argsort, np.unique first indices and counts), whose argsort is unstable (member order
 * within a structure is not part of the contract).  *ms = device time (may be NULL). */
int ahop_group_members_device(const int32_t *d_node_of_part, int64_t n,
                              const int32_t *d_node_to_struct, int32_t nnodes, int32_t nstruct,
                              int32_t *d_struct_of_part, int32_t *d_pids, int64_t *d_offsets,
                              float *ms);

/* ---- batched per-halo property reductions ------------------------------------------- */

/* Row layout of the output (one row of AHOP_HALO_NFIELDS doubles per halo):
 * m, px, py, pz, vx, vy, vz, sigma, Lx, Ly, Lz, r, Ixx, Ixy, Ixz, Iyy, Iyz, Izz,
 * sha, shb, shc, ek, ep, et, rvir, mvir, cvel, tvir, spin, rho_0, r_c
 * (rho_0, r_c: the truncated isothermal sphere of compute_halo_profile_1b41,
 * src/compute_halo_props.py:1232-1238) */
#define AHOP_HALO_NFIELDS 31

typedef struct {
    double lbox;            /* Lbox_pt, periodic box side in the members' length unit */
    double hub_pt;          /* Hub_pt */
    double gravconst;       /* gravconst, 430.1 in (km/s)^2 Mpc / 1e11 Msun */
    double vir_overdens;    /* vir_overdens of virial_12 */
    double rho_mean;        /* rho_mean of virial_12 */
    int32_t cdm;            /* 1: centre = member nearest the centre of mass; 0: densest member */
    int32_t nshell;         /* ellipsoidal shells (ttab = 1000 in the reference) */
    int32_t direct_pe_nmax; /* direct pair-sum potential energy below this count (1000) */
    int32_t tile;           /* members per work tile; <= 0 lets the autotuner choose */
    int32_t workgroup;      /* 128 or 256; <= 0 lets the autotuner choose */
    int32_t sc;             /* SC of refine_centers_1b (:1555-1596): shrinking-sphere recentring */
    int32_t dpmmc;          /* DPMMC: multiscale-grid recentring before SC (needs d_density) */
    double dcell_min;       /* dcell_min: smallest cell / sphere radius, > 0 when SC or DPMMC */
    double eps_sc;          /* eps_SC: sphere shrink fraction, in (0, 1) when SC */
    int32_t light;          /* 1: skip the shell tables, energies and virial scan (ep, et,
                               rvir, mvir, cvel, tvir, spin, rho_0, r_c read NaN) */
    /* switches of the virial potential profile (potential_profile_1b40,
     * src/compute_halo_props.py:1065-1123), 0 = the reference's default */
    int32_t virial_pe_profile;       /* 0 'shell', 1 'baseline' (-3/5 G M(<r)^2 RF(a^2,b^2,c^2)) */
    int32_t virial_pe_norm_policy;   /* 0 'profile_outer_for_large', 1 'legacy' */
    int32_t virial_shell_self_mode;  /* 0 'count', 1 'half_shell' */
    /* ANG_MOM_OF_R: per-shell mass and angular momentum of det_ang_momentum_per_shell_1b5
     * (src/compute_halo_props.py:1970-2013, called at :1257 for halos with a finite rvir).
     * ang_mom_nshells 0 turns it off (the reference's nshells is 100, halo_defs.py:595);
     * d_ang_mom (nhalos x 4 x ang_mom_nshells: per shell m, then Lx, Ly, Lz) and d_ang_axes
     * (nhalos x 12: amax, bmax, cmax, then the principal axes v row-major) are device arrays;
     * rows of halos whose rvir is not finite read NaN.  Needs light = 0. */
    int32_t ang_mom_nshells;
    double *d_ang_mom;
    double *d_ang_axes;
} ahop_halo_params;
/* lbox = +inf turns the periodic wrap off everywhere (FlagPeriod = 0, :862 and :889). */

typedef struct {
    int32_t tile, workgroup, trials;
    float ms;               /* device time of the final run, all passes */
    float ms_direct_pe;     /* of which the pair-sum potential energies (det_halo_energies_1b3) */
    float ms_virial;        /* and the shell tables, profile and virial scan (det_vir_props_1b4) */
} ahop_halo_info;

/* Properties of every halo of a CSR member layout in one call.  offsets: host array of
 * nhalos+1 member offsets; d_pos/d_vel (M,3) row-major, d_mass (M), d_density (M, may be
 * NULL when cdm=1) and d_out (nhalos x AHOP_HALO_NFIELDS) are device pointers.
 * Replaces the per-halo fork-Pool loop of src/compute_halo_props.py:487-534 over
 * det_mass_17 (:754), det_center_18 (:901), refine_centers_1b (:1555, total centre; DPMMC
 * :1447 and SC :1498), compute_ang_mom_19 (:783), r_halos_1a (:828),
 * det_vir_1b (:1241: det_main_axis_1b2, det_halo_energies_1b3, det_vir_props_1b4) and
 * compute_spin_parameter_1c (:1599). */
int ahop_halo_props_device(int64_t nhalos, const int64_t *offsets, const double *d_pos,
                           const double *d_vel, const double *d_mass, const double *d_density,
                           const ahop_halo_params *params, double *d_out, ahop_halo_info *info);

/* ---- stellar properties (mode 'full', stellar payload) ----------------------------- */

/* Columns of d_out (nhalos rows):
 * ndm, nstar, mdm, m*, mcontam, sigma_dm, sigma*, px*, py*, pz*, Lx*, Ly*, Lz*, r*, r50, r90,
 * age, metal, SFR, SFR_r50, SFR_r90, SFR10, SFR10_r50, SFR10_r90, vrot, sig3d, sigcyl,
 * vrot_r50, sig3d_r50, sigcyl_r50, vrot_r90, sig3d_r90, sigcyl_r90 */
#define AHOP_STELLAR_NFIELDS 33

/* Stellar properties of every halo of a CSR member layout: det_mass_17 (:765-780),
 * det_center_18 (:963-966), the stellar parts of refine_centers_1b (:1569-1593),
 * compute_ang_mom_19 (:806-822) and r_halos_1a (:842-849), and compute_stellar_1d
 * (:1635-1751), src/compute_halo_props.py.  d_chem (M x nchem, nchem <= 9, may be NULL with
 * nchem 0): per-member abundances in CHEM_ELEMENTS order, whose mass-weighted means over each
 * halo's stars go to d_chem_out (nhalos x nchem, NaN without stars or valid values).  d_is_star (M,
 * uint8: the reference's indexps >= ndm), d_age (Gyr) and d_metal (M, read for stars), d_density
 * (M, needed when any member is a star), d_props the ahop_halo_props_device rows (the mean
 * velocity); params: lbox and the SC / DPMMC options (cdm and light ignored); massp for
 * mcontam.  Halos without stars keep the values of the reference's clear_halo (NaN, SFRs 0). */
int ahop_halo_stellar_device(int64_t nhalos, const int64_t *offsets, const double *d_pos,
                             const double *d_vel, const double *d_mass, const double *d_density,
                             const uint8_t *d_is_star, const double *d_age,
                             const double *d_metal, const double *d_chem, int32_t nchem,
                             const double *d_props, const ahop_halo_params *params,
                             double massp, double *d_out, double *d_chem_out, float *ms);

/* ---- extended profiles (compute_extended_profiles_1e, src/compute_halo_props.py:1819-1926),
 * halos whose members are all dark matter --------------------------------------------- */

/* Columns of d_prof (nhalos rows) */
#define AHOP_PROFILE_NFIELDS 6   /* vmaxcir, rmaxcir, cNFW, cNFWerr, inslope, inslopeerr */

/* Step 1: member distances from each halo's centre (d_props: the nhalos x AHOP_HALO_NFIELDS
 * rows of ahop_halo_props_device, whose centre and rvir are used), sorted per halo into
 * d_sdist / d_smass (M each), their cumulative mass from each halo's first non-zero distance
 * into d_scum (M), the circular-velocity peak (:1827-1840) into d_prof columns 0-1
 * (columns 2-5 set to NaN), and per halo into d_aux (4 doubles): slot of the first non-zero
 * distance within the halo, members with 0 < d <= rvir, and the first and last such distance.
 * offsets: host, nhalos+1. */
int ahop_halo_profile_radii_device(int64_t nhalos, const int64_t *offsets, const double *d_pos,
                                   const double *d_mass, const double *d_props, double lbox,
                                   double gravconst, double *d_sdist, double *d_smass,
                                   double *d_scum, double *d_aux, double *d_prof, float *ms);

/* Step 2: the NFW fit of the binned log-density (:1861-1891) and the inner slope
 * (:1893-1926) into d_prof columns 2-5, for the halos given bin edges: edge_offsets (host,
 * nhalos+1) index d_edges; a halo with no edges is skipped (the reference's early returns:
 * fewer than 8 non-zero distances, rvir not > 0, fewer than 8 members inside rvir), else it
 * has nbins+1 edges, nbins = max(8, min(80, int(sqrt(members inside rvir)))), computed by
 * the caller as np.logspace(log10(d_first), log10(d_last), nbins+1). */
int ahop_halo_profile_fits_device(int64_t nhalos, const int64_t *offsets, const double *d_sdist,
                                  const double *d_smass, const double *d_scum,
                                  const double *d_props, const double *d_aux,
                                  const int64_t *edge_offsets, const double *d_edges,
                                  double *d_prof, float *ms);

/* ---- RAMSES particle files (host, one OpenMP thread per file) ------------------------ */

typedef struct {
    int64_t ndm, nstar, nselected;   /* per file: dark matter, stars, selected by family */
} ahop_ramses_counts;

/* The counting sweep of read_ramses_new_101 (src/input_output.py:842-890) over
 * repository/part_<nchar>.out<icpu:05d>, icpu = 1..ncpu: per-file counts into per_file
 * (ncpu entries).  rver 3 (Ra3: families from birth time and id sign) or 4 (Ra4: family
 * byte); family 0 all, 1 dm, 2 star; stellar: the stellar payload flag; nthreads <= 0: all. */
int ahop_ramses_count(const char *repository, const char *nchar, int32_t ncpu, int32_t rver,
                      int32_t family, int32_t stellar, int32_t nthreads,
                      ahop_ramses_counts *per_file);

/* The reading sweep (_read_ramses_new_1010, :624-760): selected particles scattered to their
 * dense-id slots (stars after ndm_total dark-matter slots with family all): pos (nselected x
 * 3, shifted by -0.5), vel (x 3, times scale_l / scale_t * 1e-5), mass; source_pid (int64,
 * families dm / star, may be NULL); birth and metal (nstar_total star slots, may be NULL):
 * the raw birth epoch, which the caller turns into an age with its cosmology
 * (snapshot_age - epoch_to_age(birth), :748-749).  Ids outside the dense range or repeated
 * within a file are errors, as in the reference. */
int ahop_ramses_read(const char *repository, const char *nchar, int32_t ncpu, int32_t rver,
                     int32_t family, int32_t stellar, int32_t nthreads, int64_t ndm_total,
                     int64_t nselected_total, int64_t nstar_total, double scale_l,
                     double scale_t, double *pos, double *vel, double *mass,
                     int64_t *source_pid, double *birth, double *metal, double *m0,
                     double *chem, ahop_ramses_counts *per_file);

/* The Ra4 field descriptor of an output (repository/part_file_descriptor.txt, parsed as
 * _parse_ra4_part_descriptor and _ra4_stellar_descriptor do, src/input_output.py:530-576):
 * *present = 1 when the file exists, *has_m0 = 1 when it lists initial_mass (then
 * ahop_ramses_read fills m0, nstar_total slots), and *chem_mask bit e when it lists
 * chem_<element e> for the elements H, O, Fe, Mg, C, N, Si, S, D (src/chem_species.py:1; then
 * ahop_ramses_read fills chem, nstar_total x 9 row-major, column e).  With a descriptor the
 * stellar records after the tag are read where it places them and the other fields skipped
 * (_read_ra4_stellar_record_block, :580-620); without one, birth times and metallicities are
 * the two records after the tag.  The reference's errors: initial_mass without metallicity,
 * or before it. */
int ahop_ramses_descriptor(const char *repository, int32_t *present, int32_t *has_m0,
                           int32_t *chem_mask);

/* Host structure-tree stage alone (create_group_tree f90:885-983) on host inputs, for
 * callers that computed the hot-path stages elsewhere.  Outputs: liste_parts (npart),
 * *nnodes, and the tables, which must have room for 2*ngroups nodes (nnodes <= 2*ngroups,
 * f90:906): real_table 8*2*ngroups doubles, integer_table 6*2*ngroups int32.  The node
 * creation follows the optimize_nodes switch of the last ahop_sync_others call (default
 * off), as the reference's module variable does (f90:954-958). */
int ahop_create_group_tree(int64_t npart, const double *pos, const double *mass,
                           const double *density, const int32_t *gid, int32_t ngroups,
                           const double *densityg, const int32_t *sad_offset,
                           const int32_t *sad_isad, const double *sad_rho, double rho_threshold,
                           double fudge, double alphap, double epsilon, int32_t nmembthresh,
                           double xlong, int32_t *liste_parts, int32_t *nnodes,
                           double *real_table, int32_t *integer_table);

/* Counters of the last host structure-tree stage (ahop_compute_adaptahop or
 * ahop_create_group_tree; create_group_tree f90:885-983 and omp_dev_create_nodes_task
 * :1245-1628 run it with OpenMP tasks): out[0] levels whose colour components were treated
 * in parallel (an OpenMP taskloop), out[1] the components treated there, out[2] child subtrees
 * deferred as tasks, out[3] nodes.  out_ms (may be NULL): milliseconds of linking the group
 * members, the first colouring, node creation, the deterministic reordering and the tables. */
int ahop_tree_stats(int64_t *out /* [4] */, double *out_ms /* [5] */);

#ifdef __cplusplus
}
#endif
#endif
