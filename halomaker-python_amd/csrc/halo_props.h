// halo_props.h -- batched per-halo property reductions (see halo_props.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ahop {

// Output row layout (one row of HP_NFIELDS doubles per halo); names follow halo_dtype
// (src/halo_defs.py:302-334) and the inertia tensor of det_inertial_tensor_1b20.
enum HaloField {
    HP_M = 0, HP_PX, HP_PY, HP_PZ, HP_VX, HP_VY, HP_VZ, HP_SIGMA, HP_LX, HP_LY, HP_LZ, HP_R,
    HP_IXX, HP_IXY, HP_IXZ, HP_IYY, HP_IYZ, HP_IZZ, HP_SHA, HP_SHB, HP_SHC, HP_EK, HP_EP, HP_ET,
    HP_RVIR, HP_MVIR, HP_CVEL, HP_TVIR, HP_SPIN, HP_RHO0, HP_RC, HP_NFIELDS
};

struct HaloParamsDev {
    double lbox = 1.0;          // Lbox_pt: periodic box side in the members' units
    double hub_pt = 70.3;       // Hub_pt (km/s/Mpc)
    double gravconst = 430.1;   // gravconst (halo_defs.py:482)
    double vir_overdens = 340.0;
    double rho_mean = 1.0;
    int cdm = 1;                // 1: member nearest the COM (det_center_18 cdm branch); 0: densest
    int nshell = 1000;          // ellipsoidal shells of tab_props_inside_1b40 (ttab)
    int direct_pe_nmax = 1000;  // pair-sum potential energy below this member count
    int sc = 0;                 // SC: shrinking-sphere recentring (refine_centers_1b)
    int dpmmc = 0;              // DPMMC: densest member of the most massive cell, then SC if on
    double dcell_min = 0.0;     // smallest cell / sphere radius of the recentring
    double eps_sc = 0.0;        // sphere radius shrink fraction of SC
    int light = 0;              // 1: mass, centre, mean velocity, sigma, L, r, I, axes, ek only
                                // (ep, et, rvir, mvir, cvel, tvir, spin, rho_0, r_c left NaN)
    // the virial potential profile's switches (potential_profile_1b40 :1065-1123); 0 is the
    // reference's default of each
    int pe_baseline = 0;        // virial_pe_profile: 0 'shell', 1 'baseline' (-3/5 G M(<r)^2 RF)
    int pe_norm_legacy = 0;     // virial_pe_norm_policy: 0 'profile_outer_for_large', 1 'legacy'
    int shell_half = 0;         // virial_shell_self_mode: 0 'count', 1 'half_shell'
    // ANG_MOM_OF_R (det_ang_momentum_per_shell_1b5): shells of the per-shell angular momentum
    // (0: off), and its device outputs (nhalos x 4 x shells; nhalos x 12)
    int angmom_nshells = 0;
    double *angmom = nullptr;
    double *angmom_axes = nullptr;
};

struct HaloTuning {
    int tile = 0, bs = 0, trials = 0;
    float ms = 0;   // device time of the final run, all passes
    // of which the pair-sum potential (k_direct_pe) and the shell tables / virial scan
    // (k_hist_fold, k_virial_lanes, k_virial_serial); 0 in light mode
    float sub_ms[2] = {0, 0};
};

// Stellar properties (halo_stellar.hip): is_star / age / metal per member, chem (M x nchem,
// may be null with nchem 0), props the nhalos x HP_NFIELDS rows of halo_props_device (mean
// velocity), out nhalos x 33, chem_out nhalos x nchem.
float halo_stellar_device(const int64_t *h_off, int64_t nh, const double *pos, const double *vel,
                          const double *mass, const double *dens, const uint8_t *is_star,
                          const double *age, const double *metal, const double *chem,
                          int nchem, const double *props, const HaloParamsDev &P, double massp,
                          double *out, double *chem_out, hipStream_t s);

// pos/vel (M,3), mass (M), dens (M, may be null when cdm) are device pointers in CSR order;
// h_off (nhalos+1) host offsets; d_out device (nhalos x HP_NFIELDS).  tile/bs <= 0 autotune.
HaloTuning halo_props_device(const int64_t *h_off, int64_t nhalos, const double *pos,
                             const double *vel, const double *mass, const double *dens,
                             const HaloParamsDev &P, double *d_out, int tile, int bs,
                             hipStream_t s);
}  // namespace ahop
