// ahop_common.h -- shared definitions of the MI355X AdaptaHOP hot path.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

namespace ahop {

// Hot-path parameters (the subset of the reference's module state that the
// kNN/density/hop/saddle stages read: compute_adaptahop.f90:682-881, :2433-2962).
struct Params {
    int nvoisins = 32;          // k of the kNN (self excluded)
    int nhop = 16;              // neighbours scanned for the hop target / saddles
    double rho_threshold = 80.0;
    double xlong = 1.0, ylong = 1.0, zlong = 1.0;
    int mixed = 1;              // 1: FP32 box-relative filter + FP64 re-rank, 0: FP64 throughout
    int bucket = 64;            // particles per leaf bucket of the device tree (<= 64)
};

// Device tree built over Morton-sorted particles (binary radix tree, leaves = particles).
// Internal node i in [0, n-2] covers sorted range [first[i], last[i]].
// A child reference c >= 0 is internal; c < 0 encodes particle ~c.
struct DeviceTree {
    int64_t n = 0;
    uint64_t *key = nullptr, *key_tmp = nullptr;
    int32_t *perm = nullptr, *iota = nullptr;   // sorted slot -> original index
    double4 *spos = nullptr;                    // sorted (x, y, z, mass)
    int32_t *first = nullptr, *last = nullptr;
    int32_t *left = nullptr, *right = nullptr;
    int32_t *parent = nullptr, *leaf_parent = nullptr;
    int32_t *flag = nullptr;
    double *box = nullptr;                      // 6 doubles per internal node: lo xyz, hi xyz
    float4 *fbox = nullptr;                     // FP32 centre / half-widths (2 per node)
    // per internal node, what a visit of it reads, in one place: (left, right, first, last),
    // and the FP32 boxes of its two children with their slot ranges: child k's
    // (cx, cy, cz, hx) and (hy, hz, first, last) (ints as float bits), k = 1, 2
    int4 *nrec = nullptr;
    float4 *cbox = nullptr;
};

// Per-stage device timings of the last run (milliseconds, HIP events on the run's stream).
struct StageTimes {
    float sort = 0, tree = 0, knn = 0, fallback = 0, hop = 0, saddle = 0, total = 0;
};

// Results of the hot path in original particle order, resident on the device.
struct DeviceResults {
    double *density = nullptr;        // [n]
    int32_t *gid = nullptr;           // [n] 1-based group id, 0 = inactive
    int32_t *hop = nullptr;           // [n] 0-based original index of hop target, -1 inactive
    int32_t *nbr = nullptr;           // [n*nhop] sorted-slot neighbour ids (internal)
    int64_t nactive = 0;
    int32_t ngroups = 0;
    int64_t nfallback = 0;
    // saddle CSR (0-based offsets, 1-based group ids)
    int32_t *sad_offset = nullptr;    // [ngroups+1]
    int32_t *sad_isad = nullptr;
    double *sad_rho = nullptr;
    double *densityg = nullptr;       // [ngroups]
    int64_t nsaddle = 0;
    // work counters of the main kNN launch: nodes visited, buckets scanned, candidate
    // distances per lane, heap-insertion steps, heap insertions, waves, and of the selection
    // kernel the largest per-lane log, the lanes whose log overflowed and the lanes whose
    // FP32 order of the k nearest differed from the FP64 order (both re-run in FP64)
    unsigned long long knn_stats[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

// Growable device allocation owned by an Engine.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void *ensure(size_t nbytes);
    void release();
    ~DevBuf() { release(); }
};

struct KnnArgs;   // kNN kernel arguments (ahop_kernels.hip)

class Engine {
  public:
    Engine();
    ~Engine();
    // Full hot path on device-resident inputs: pos (n,3) row-major doubles, mass (n).
    // Fills results (device pointers valid until the next run or close).
    void run(const double *d_pos, const double *d_mass, int64_t n, const Params &p, bool saddle);
    // kNN + density only, with every particle's k neighbours (original 0-based ids, ascending
    // distance) written to d_nbr (n*k) and the density to d_rho (n), both in original order
    // d_order (n, may be null) receives the Morton order (sorted slot -> original index)
    void knn_lists(const double *d_pos, const double *d_mass, int64_t n, const Params &p,
                   double *d_rho, int32_t *d_nbr, int32_t *d_order);
    const DeviceResults &results() const { return res_; }
    const StageTimes &times() const { return times_; }
    hipStream_t stream() const { return stream_; }
    void release_all();

  private:
    void build_tree(const double *d_pos, const double *d_mass, int64_t n, const Params &p);
    void knn_density(const Params &p);
    void launch_knn_sel(const KnnArgs &a, unsigned nw, hipStream_t s);
    void hop_groups(const Params &p);
    void saddles(const Params &p);

    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[8];
    DeviceTree t_;
    DeviceResults res_;
    StageTimes times_;
    // buffers
    DevBuf b_key, b_key_tmp, b_perm, b_iota, b_spos, b_first, b_last, b_left, b_right, b_parent,
        b_leaf_parent, b_flag, b_box, b_temp, b_rho_s, b_nbr, b_hop_s, b_gid_s, b_flag_o,
        b_rank_o, b_density, b_gid, b_hop, b_counter, b_fallback, b_emit_cnt, b_emit_off,
        b_ekey, b_eval, b_ekey2, b_eval2, b_ukey, b_uval, b_keep, b_keep_off, b_skey, b_sval,
        b_sad_off, b_sad_isad, b_sad_rho, b_sad_cnt, b_densityg, b_misc, b_roots, b_fbox, b_fb_sorted,
        b_nrec, b_cbox, b_klog, b_wctr;
    int sel_cus_ = 0;   // compute units of the device (persistent grid of the selection kernel)
    int64_t nmax_ = 0;
};

const char *last_error();
void set_error(const std::string &msg);

// Adjoint of the SPH density with frozen neighbour rows (sph_adjoint.hip): all pointers are
// device arrays in original order (pos (n,3), nbr (n,K)); returns the device milliseconds.
// d_order (n, may be null): a permutation of the particles that keeps neighbours close (the
// Morton order of the forward); with it, each workgroup takes consecutive positions of that
// order and adds its neighbours' terms into a window of the order in LDS first.
// Particles grouped by label: labels = node_to_struct[node - 1] for node > 0 (0 stays 0; the
// node ids themselves when node_to_struct is null, nlab = nnodes).  Writes the labels
// (d_labels_out, may be null), the particle indices sorted by label, ascending within a label
// (d_pids, n), and CSR offsets (d_offsets, nlab + 2: label l holds d_pids[off[l], off[l+1])).
// Returns the device time in ms.
float group_members(const int32_t *d_node, int64_t n, const int32_t *d_node_to_struct,
                    int32_t nnodes, int32_t nlab, int32_t *d_labels_out, int32_t *d_pids,
                    int64_t *d_offsets, hipStream_t s);

// Extended profiles (halo_profiles.hip).  radii: distances from each halo's centre sorted per
// halo (d_sdist, d_smass: M; d_scum: M, the cumulative mass from each halo's first
// non-zero distance), per halo aux (first non-zero slot, members inside rvir, first
// and last such distance) and vmaxcir, rmaxcir in d_prof (nh x 6; the fit columns NaN).
// fits: with the caller's bin edges (h_eoff: host nh+1 offsets into d_edges; 0 edges = no
// fits) the NFW and inner-slope fits.  props: nh x 29 halo property rows.  Device ms.
float halo_profile_radii(const int64_t *h_off, int64_t nh, const double *d_pos,
                         const double *d_mass, const double *d_props, double L, double G,
                         double *d_sdist, double *d_smass, double *d_scum, double *d_aux,
                         double *d_prof, hipStream_t s);
float halo_profile_fits(const int64_t *h_off, int64_t nh, const double *d_sdist,
                        const double *d_smass, const double *d_scum, const double *d_props,
                        const double *d_aux,
                        const int64_t *h_eoff, const double *d_edges, double *d_prof,
                        hipStream_t s);

float sph_density_adjoint(const double *d_pos, const double *d_mass, const int32_t *d_nbr,
                          const int32_t *d_order, const double *d_grad_rho, int64_t n, int K,
                          double Lx, double Ly, double Lz, double *d_gpos, double *d_gmass,
                          hipStream_t s);
// The same without allocating: its scratch (an error counter and the order's inverse) is a
// fixed device array of the library, n <= 4194304; calls overlapping on one GPU race.
void sph_density_adjoint_noalloc(const double *d_pos, const double *d_mass,
                                 const int32_t *d_nbr, const int32_t *d_order,
                                 const double *d_grad_rho, int64_t n, int K, double Lx, double Ly,
                                 double Lz, double *d_gpos, double *d_gmass, hipStream_t s);

}  // namespace ahop

#define AHOP_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) +     \
                                     " at " + __FILE__ + ":" + std::to_string(__LINE__));     \
    } while (0)
