// ahop_capi.cpp -- C ABI (include/amdahop.h) over the device engine and the host
// structure-tree stage.  Holds the process-wide module state that the reference keeps
// in Fortran module variables (fhalo_defs / neikdtree, compute_adaptahop.f90:12-407).
#include "../../include/amdahop.h"
#include "ahop_common.h"
#include "create_nodes.h"
#include "halo_props.h"
#include "ramses_reader.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using namespace ahop;

namespace {

// Host output buffer in pinned (page-locked) memory: device-to-host copies into it run at
// link speed instead of through pageable staging.  The capacity is kept across calls (pinning
// is costly) and released only at process end; clear() only resets the size.
template <typename T>
struct PinnedVec {
    T *p = nullptr;
    size_t n = 0, cap = 0;
    void resize(size_t m)
    {
        if (m > cap) {
            if (p) AHOP_HIP(hipHostFree(p));
            p = nullptr;
            cap = 0;
            AHOP_HIP(hipHostMalloc((void **)&p, sizeof(T) * m, hipHostMallocDefault));
            cap = m;
        }
        n = m;
    }
    void clear() { n = 0; }
    T *data() { return p; }
    const T *data() const { return p; }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
    // no destructor: the HIP runtime may already be gone when static state is destroyed
};

struct ModuleState {
    // sync_from_init_adaptahop (f90:591-641)
    int64_t npart = 0;
    int32_t nmembthresh = 0, nmembers = 0;
    double omegal = 0, omega_lambda_f = 0, omega0 = 0, omega_f = 0, aexp_max = 0, af = 0;
    double hubble = 0, h_f = 0, boxsize2 = 1, lf = 1;
    double xlong = 1, ylong = 1, zlong = 1, xlongs2 = 0.5, ylongs2 = 0.5, zlongs2 = 0.5;
    double hub_pt = 0, aexp = 0, pos_shift[3] = {0, 0, 0};
    bool fsub = false;
    // sync_from_change_pos (f90:576-589)
    int32_t nusedpart = 0;
    double epsilon = 0, fudgepsilon = 0;
    // sync_others (f90:643-678)
    bool verbose = false, megaverbose = false;
    int32_t nbpes = 1;
    double rho_threshold = 80, massp = 0, boxsize = 0;
    int32_t nhop = 16, nvoisins = 32;
    double fudge = 4, alphap = 1;
    std::string method = "MSM";
    int32_t nlevelmax = 30;
    bool optimize_nodes = false;
    // extensions
    int32_t mixed = 1, bucket = 64;
    // outputs (allocated by compute_adaptahop, freed by close)
    bool allocated = false;
    std::vector<int32_t> liste_parts;
    PinnedVec<double> density;
    int32_t nnodes = 0;
    std::vector<double> real_table;
    std::vector<int32_t> integer_table;
    PinnedVec<int32_t> gid, hop;
    std::vector<int32_t> sad_offset, sad_isad;
    std::vector<double> sad_rho, densityg;
    int32_t ngroups = 0;
    // last device run (ahop_run_device)
    std::vector<int32_t> dev_sad_offset, dev_sad_isad;
    std::vector<double> dev_sad_rho;
    int64_t dev_n = 0;   // particle count of the last ahop_run_device (0: none)
    // counters and stage times of the last host structure-tree stage
    int64_t tree_stats[4] = {0, 0, 0, 0};
    double tree_ms[5] = {0, 0, 0, 0, 0};
    // compute_adaptahop's device copy of its inputs (positions, then masses) and the pinned
    // staging buffers of their upload: kept across calls, released at process end
    void *dev_in = nullptr;
    size_t dev_in_cap = 0;
    PinnedVec<char> stage[2];
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

void keep_tree_stats(ModuleState &s, const TreeOutputs &to)
{
    s.tree_stats[0] = to.taskloop_levels;
    s.tree_stats[1] = to.taskloop_components;
    s.tree_stats[2] = to.subtree_tasks;
    s.tree_stats[3] = to.nnodes;
    s.tree_ms[0] = to.ms_link;
    s.tree_ms[1] = to.ms_colour;
    s.tree_ms[2] = to.ms_nodes;
    s.tree_ms[3] = to.ms_reorder;
    s.tree_ms[4] = to.ms_output;
}

ModuleState &S()
{
    static ModuleState s;
    return s;
}

// Host-to-device copy of a caller's pageable array: host threads copy 32 MiB chunks into two
// pinned staging buffers in turn while the DMA engine moves the previous chunk (a plain
// hipMemcpy from pageable memory moved the 0.54 GB of a 16.8M-particle box at about 27 GB/s).
void upload(ModuleState &s, void *dst, const void *src, size_t bytes, hipStream_t st)
{
    constexpr size_t CH = size_t(32) << 20;
    for (int b = 0; b < 2; ++b) {
        s.stage[b].resize(CH);
        if (!s.stage_ev[b])
            AHOP_HIP(hipEventCreateWithFlags(&s.stage_ev[b], hipEventDisableTiming));
    }
    const int nt = std::max(1, std::min(omp_get_max_threads(), 16));
    bool pending[2] = {false, false};
    int b = 0;
    for (size_t off = 0; off < bytes; off += CH, b ^= 1) {
        const size_t m = std::min(CH, bytes - off);
        if (pending[b]) AHOP_HIP(hipEventSynchronize(s.stage_ev[b]));   // buffer b is free
        char *stg = s.stage[b].data();
        const char *from = static_cast<const char *>(src) + off;
#pragma omp parallel for num_threads(nt) schedule(static, 1)
        for (int t = 0; t < nt; ++t) {
            const size_t a = m * t / nt, e = m * (t + 1) / nt;
            std::memcpy(stg + a, from + a, e - a);
        }
        AHOP_HIP(hipMemcpyAsync(static_cast<char *>(dst) + off, stg, m, hipMemcpyHostToDevice,
                                st));
        AHOP_HIP(hipEventRecord(s.stage_ev[b], st));
        pending[b] = true;
    }
    AHOP_HIP(hipStreamSynchronize(st));
}

Engine &engine()
{
    static std::unique_ptr<Engine> e;
    if (!e) e.reset(new Engine());
    return *e;
}

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        set_error("");
        return 0;
    } catch (const std::exception &ex) {
        set_error(ex.what());
        return 1;
    } catch (...) {
        set_error("unknown error");
        return 1;
    }
}

void free_outputs(ModuleState &s)
{
    s.allocated = false;
    // liste_parts keeps its buffer for the next call (refilled there; unreadable meanwhile)
    s.density.clear();
    s.real_table.clear();
    s.integer_table.clear();
    s.nnodes = 0;
    s.gid.clear();
    s.hop.clear();
    s.sad_offset.clear();
    s.sad_isad.clear();
    s.sad_rho.clear();
    s.densityg.clear();
    s.ngroups = 0;
}

Params hot_params(const ModuleState &s)
{
    Params p;
    p.nvoisins = s.nvoisins;
    p.nhop = s.nhop;
    p.rho_threshold = s.rho_threshold;
    p.xlong = s.xlong;
    p.ylong = s.ylong;
    p.zlong = s.zlong;
    p.mixed = s.mixed;
    p.bucket = s.bucket;
    return p;
}

}  // namespace

extern "C" {

int ahop_sync_from_change_pos(int64_t npart, int32_t nusedpart, double epsilon, double fudgepsilon,
                              double xlong, double boxsize2)
{
    return guarded([&] {
        auto &s = S();
        s.npart = npart;
        s.nusedpart = nusedpart;
        s.epsilon = epsilon;
        s.fudgepsilon = fudgepsilon;
        s.xlong = xlong;
        s.boxsize2 = boxsize2;
    });
}

int ahop_sync_from_init_adaptahop(int64_t npart, int32_t nmembthresh, int32_t nmembers,
                                  double omegal, double omega_lambda_f, double omega0,
                                  double omega_f, double aexp_max, double af, double hubble,
                                  double h_f, double boxsize2, double lf, double xlong,
                                  double ylong, double zlong, double xlongs2, double ylongs2,
                                  double zlongs2, double hub_pt, double aexp,
                                  const double *pos_shift)
{
    return guarded([&] {
        auto &s = S();
        s.npart = npart;
        s.nmembthresh = nmembthresh;
        s.nmembers = nmembers;
        s.omegal = omegal;
        s.omega_lambda_f = omega_lambda_f;
        s.omega0 = omega0;
        s.omega_f = omega_f;
        s.aexp_max = aexp_max;
        s.af = af;
        s.hubble = hubble;
        s.h_f = h_f;
        s.boxsize2 = boxsize2;
        s.lf = lf;
        s.xlong = xlong;
        s.ylong = ylong;
        s.zlong = zlong;
        s.xlongs2 = xlongs2;
        s.ylongs2 = ylongs2;
        s.zlongs2 = zlongs2;
        s.hub_pt = hub_pt;
        s.aexp = aexp;
        if (pos_shift)
            for (int d = 0; d < 3; ++d) s.pos_shift[d] = pos_shift[d];
        s.fsub = true;
    });
}

int ahop_sync_others(int32_t verbose, int32_t megaverbose, int32_t npart, int32_t nbpes,
                     double rho_threshold, double massp, double boxsize, int32_t nhop,
                     int32_t nvoisins, double fudge, double alphap, const char *method,
                     int32_t method_len, int32_t nlevelmax, int32_t optimize_nodes)
{
    return guarded([&] {
        auto &s = S();
        s.verbose = verbose != 0;
        s.megaverbose = megaverbose != 0;
        s.npart = npart;
        s.nbpes = nbpes;
        s.rho_threshold = rho_threshold;
        s.massp = massp;
        s.boxsize = boxsize;
        s.nhop = nhop;
        s.nvoisins = nvoisins;
        s.fudge = fudge;
        s.alphap = alphap;
        // method is character(len=3) in fhalo_defs: keep at most 3 characters (f90:669)
        s.method = method ? std::string(method, method_len < 3 ? (method_len < 0 ? 0 : method_len) : 3) : "";
        s.nlevelmax = nlevelmax;
        s.optimize_nodes = optimize_nodes != 0;
    });
}

int ahop_close(void)
{
    return guarded([&] { free_outputs(S()); });
}

int ahop_compute_adaptahop(const double *pos_in, int64_t pos_rows, int64_t pos_cols,
                           const double *mass_in, int64_t mass_len)
{
    return guarded([&] {
        auto &s = S();
        free_outputs(s);                           // f90:501, close() first
        if (pos_rows != 3) throw std::runtime_error("pos_in should have shape (3,npart)");
        const int64_t n = s.npart;
        if (n <= 0) throw std::runtime_error("npart must be set (sync_* calls) before compute_adaptahop");
        if (pos_cols != n || mass_len != n)
            throw std::runtime_error("pos_in/mass_in sizes do not match npart=" + std::to_string(n));
        if (s.nusedpart < n) throw std::runtime_error("nusedpart must be >= npart");
        if (!pos_in || !mass_in) throw std::runtime_error("null input array");
        auto t0 = std::chrono::steady_clock::now();
        s.dev_n = 0;   // the engine's results now belong to this call
        Engine &e = engine();
        Params p = hot_params(s);
        const size_t need = sizeof(double) * 4 * n;
        if (need > s.dev_in_cap) {
            if (s.dev_in) AHOP_HIP(hipFree(s.dev_in));
            s.dev_in = nullptr;
            s.dev_in_cap = 0;
            AHOP_HIP(hipMalloc(&s.dev_in, need));
            s.dev_in_cap = need;
        }
        double *dpos = static_cast<double *>(s.dev_in), *dmass = dpos + 3 * n;
        upload(s, dpos, pos_in, sizeof(double) * 3 * n, e.stream());
        upload(s, dmass, mass_in, sizeof(double) * n, e.stream());
        e.run(dpos, dmass, n, p, true);
        const DeviceResults &r = e.results();
        s.density.resize(n);
        s.gid.resize(n);
        s.hop.resize(n);
        s.ngroups = r.ngroups;
        s.densityg.resize(r.ngroups);
        s.sad_offset.resize(r.ngroups + 1);
        s.sad_isad.resize(r.nsaddle);
        s.sad_rho.resize(r.nsaddle);
        AHOP_HIP(hipMemcpy(s.density.data(), r.density, sizeof(double) * n, hipMemcpyDeviceToHost));
        AHOP_HIP(hipMemcpy(s.gid.data(), r.gid, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        AHOP_HIP(hipMemcpy(s.hop.data(), r.hop, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        if (r.ngroups > 0)
            AHOP_HIP(hipMemcpy(s.densityg.data(), r.densityg, sizeof(double) * r.ngroups,
                               hipMemcpyDeviceToHost));
        AHOP_HIP(hipMemcpy(s.sad_offset.data(), r.sad_offset, sizeof(int32_t) * (r.ngroups + 1),
                           hipMemcpyDeviceToHost));
        if (r.nsaddle > 0) {
            AHOP_HIP(hipMemcpy(s.sad_isad.data(), r.sad_isad, sizeof(int32_t) * r.nsaddle,
                               hipMemcpyDeviceToHost));
            AHOP_HIP(hipMemcpy(s.sad_rho.data(), r.sad_rho, sizeof(double) * r.nsaddle,
                               hipMemcpyDeviceToHost));
        }
        auto t1 = std::chrono::steady_clock::now();
        TreeInputs ti;
        ti.npart = n;
        ti.pos = pos_in;
        ti.mass = mass_in;
        ti.density = s.density.data();
        ti.gid = s.gid.data();
        ti.ngroups = s.ngroups;
        ti.densityg = s.densityg.data();
        ti.sad_offset = s.sad_offset.data();
        ti.sad_isad = s.sad_isad.data();
        ti.sad_rho = s.sad_rho.data();
        TreeParams tp;
        tp.rho_threshold = s.rho_threshold;
        tp.fudge = s.fudge;
        tp.alphap = s.alphap;
        tp.epsilon = s.epsilon;
        tp.nmembthresh = s.nmembthresh;
        tp.xlong = s.xlong;
        tp.ylong = s.ylong;
        tp.zlong = s.zlong;
        tp.xlongs2 = s.xlongs2;
        tp.ylongs2 = s.ylongs2;
        tp.zlongs2 = s.zlongs2;
        tp.nthreads = s.nbpes;   // the reference runs this stage on nbPes OpenMP threads
        tp.optimize_nodes = s.optimize_nodes;
        TreeOutputs to;
        to.liste_parts = std::move(s.liste_parts);   // recycle the previous call's buffer
        create_group_tree(ti, tp, to);
        keep_tree_stats(s, to);
        auto t2 = std::chrono::steady_clock::now();
        if ((int64_t)to.liste_parts.size() == (int64_t)s.nusedpart) {
            s.liste_parts = std::move(to.liste_parts);
        } else {
            s.liste_parts.assign(s.nusedpart, 0);
            std::copy(to.liste_parts.begin(), to.liste_parts.end(), s.liste_parts.begin());
        }
        s.nnodes = to.nnodes;
        s.real_table = std::move(to.real_table);
        s.integer_table = std::move(to.integer_table);
        s.allocated = true;
        if (s.verbose) {
            const StageTimes &t = e.times();
            double host_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
            double all_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
            std::fprintf(stderr,
                         "    [amdahop] sort %.3f ms, tree %.3f ms, kNN+density %.3f ms "
                         "(FP64 re-run of %lld queries %.3f ms), hop+groups %.3f ms, saddle %.3f ms\n"
                         "    [amdahop] active %lld, groups %d, saddle entries %lld, nodes %d; "
                         "create nodes (host) %.3f ms, call total %.3f ms\n"
                         "    [amdahop] create nodes: link %.3f, colour %.3f, nodes %.3f, "
                         "reorder %.3f, tables %.3f ms\n",
                         t.sort, t.tree, t.knn, (long long)r.nfallback, t.fallback, t.hop,
                         t.saddle, (long long)r.nactive, r.ngroups, (long long)r.nsaddle, s.nnodes,
                         host_ms, all_ms, to.ms_link, to.ms_colour, to.ms_nodes, to.ms_reorder,
                         to.ms_output);
        }
    });
}

const int32_t *ahop_liste_parts(int64_t *len)
{
    auto &s = S();
    if (len) *len = s.allocated ? (int64_t)s.liste_parts.size() : 0;
    static const int32_t empty_i = 0;
    return s.allocated ? (s.liste_parts.empty() ? &empty_i : s.liste_parts.data()) : nullptr;
}

const double *ahop_density(int64_t *len)
{
    auto &s = S();
    if (len) *len = s.allocated ? (int64_t)s.density.size() : 0;
    static const double empty_d = 0;
    return s.allocated ? (s.density.empty() ? &empty_d : s.density.data()) : nullptr;
}

int32_t ahop_nnodes(void) { return S().nnodes; }

const double *ahop_real_table(int64_t *ncols)
{
    auto &s = S();
    if (ncols) *ncols = s.allocated ? s.nnodes : 0;
    static const double empty_d = 0;
    return s.allocated ? (s.real_table.empty() ? &empty_d : s.real_table.data()) : nullptr;
}

const int32_t *ahop_integer_table(int64_t *ncols)
{
    auto &s = S();
    if (ncols) *ncols = s.allocated ? s.nnodes : 0;
    static const int32_t empty_i = 0;
    return s.allocated ? (s.integer_table.empty() ? &empty_i : s.integer_table.data()) : nullptr;
}

const char *ahop_last_error(void) { return last_error(); }

int ahop_set_precision(int32_t mixed)
{
    return guarded([&] {
        if (mixed != 0 && mixed != 1) throw std::runtime_error("precision must be 0 (fp64) or 1 (mixed)");
        S().mixed = mixed;
    });
}

int ahop_set_bucket(int32_t bucket)
{
    return guarded([&] {
        if (bucket < 1 || bucket > 64) throw std::runtime_error("bucket must be in [1, 64]");
        S().bucket = bucket;
    });
}

int32_t ahop_ngroups(void) { return S().ngroups; }

int ahop_get_groups(int32_t *gid, int32_t *hop, int64_t n)
{
    return guarded([&] {
        auto &s = S();
        if (!s.allocated) throw std::runtime_error("no results: call compute_adaptahop first");
        if (n != (int64_t)s.gid.size()) throw std::runtime_error("size mismatch");
        if (gid) std::memcpy(gid, s.gid.data(), sizeof(int32_t) * n);
        if (hop) std::memcpy(hop, s.hop.data(), sizeof(int32_t) * n);
    });
}

int64_t ahop_nsaddle(void) { return (int64_t)S().sad_isad.size(); }

int ahop_get_saddle(int32_t *offset, int32_t *isad, double *rho)
{
    return guarded([&] {
        auto &s = S();
        if (!s.allocated) throw std::runtime_error("no results: call compute_adaptahop first");
        if (offset) std::memcpy(offset, s.sad_offset.data(), sizeof(int32_t) * s.sad_offset.size());
        if (isad && !s.sad_isad.empty()) std::memcpy(isad, s.sad_isad.data(), sizeof(int32_t) * s.sad_isad.size());
        if (rho && !s.sad_rho.empty()) std::memcpy(rho, s.sad_rho.data(), sizeof(double) * s.sad_rho.size());
    });
}

static Params hot_params(const ahop_hot_params *params)
{
    if (!params) throw std::runtime_error("null params");
    Params p;
    p.nvoisins = params->nvoisins;
    p.nhop = params->nhop;
    p.rho_threshold = params->rho_threshold;
    p.xlong = params->xlong;
    p.ylong = params->ylong;
    p.zlong = params->zlong;
    p.mixed = params->mixed;
    p.bucket = params->bucket;
    return p;
}

static void fill_info(ahop_run_info *info, const Engine &e, bool saddle)
{
    if (!info) return;
    const DeviceResults &r = e.results();
    const StageTimes &t = e.times();
    info->nactive = r.nactive;
    info->ngroups = r.ngroups;
    info->nfallback = r.nfallback;
    info->nsaddle = saddle ? r.nsaddle : 0;
    info->t_sort = t.sort;
    info->t_tree = t.tree;
    info->t_knn = t.knn;
    info->t_fallback = t.fallback;
    info->t_hop = t.hop;
    info->t_saddle = t.saddle;
    info->t_total = t.total;
    info->knn_nodes = (int64_t)r.knn_stats[0];
    info->knn_buckets = (int64_t)r.knn_stats[1];
    info->knn_candidates = (int64_t)r.knn_stats[2];
    info->knn_insert_steps = (int64_t)r.knn_stats[3];
    info->knn_inserts = (int64_t)r.knn_stats[4];
    info->knn_waves = (int64_t)r.knn_stats[5];
    info->knn_log_max = (int64_t)r.knn_stats[6];
    info->knn_log_overflow = (int64_t)r.knn_stats[7];
    info->knn_misorder = (int64_t)r.knn_stats[8];
}

int ahop_knn_device(const double *d_pos, const double *d_mass, int64_t n,
                    const ahop_hot_params *params, double *d_density, int32_t *d_nbr,
                    int32_t *d_order, ahop_run_info *info)
{
    return guarded([&] {
        Params p = hot_params(params);
        if (!d_pos || !d_mass || !d_density || !d_nbr) throw std::runtime_error("null array");
        S().dev_n = 0;   // the engine's buffers now belong to this call
        Engine &e = engine();
        e.knn_lists(d_pos, d_mass, n, p, d_density, d_nbr, d_order);
        fill_info(info, e, false);
    });
}

int ahop_density_adjoint_device(const double *d_pos, const double *d_mass, int64_t n, int32_t k,
                                const int32_t *d_nbr, const int32_t *d_order,
                                const double *d_grad_rho, double xlong,
                                double ylong, double zlong, double *d_grad_pos,
                                double *d_grad_mass, float *ms)
{
    return guarded([&] {
        if (!d_pos || !d_mass || !d_nbr || !d_grad_rho || !d_grad_pos || !d_grad_mass)
            throw std::runtime_error("null array");
        Engine &e = engine();
        float t = sph_density_adjoint(d_pos, d_mass, d_nbr, d_order, d_grad_rho, n, k, xlong,
                                      ylong, zlong, d_grad_pos, d_grad_mass, e.stream());
        if (ms) *ms = t;
    });
}

int ahop_density_adjoint_device_noalloc(const double *d_pos, const double *d_mass, int64_t n,
                                        int32_t k, const int32_t *d_nbr, const int32_t *d_order,
                                        const double *d_grad_rho, double xlong, double ylong,
                                        double zlong, double *d_grad_pos, double *d_grad_mass)
{
    return guarded([&] {
        if (!d_pos || !d_mass || !d_nbr || !d_grad_rho || !d_grad_pos || !d_grad_mass)
            throw std::runtime_error("null array");
        // the device's null stream: this entry point creates no stream (nor anything else);
        // the call is synchronous like the others
        sph_density_adjoint_noalloc(d_pos, d_mass, d_nbr, d_order, d_grad_rho, n, k, xlong, ylong,
                                    zlong, d_grad_pos, d_grad_mass, nullptr);
    });
}

int ahop_group_members_device(const int32_t *d_node_of_part, int64_t n,
                              const int32_t *d_node_to_struct, int32_t nnodes, int32_t nstruct,
                              int32_t *d_struct_of_part, int32_t *d_pids, int64_t *d_offsets,
                              float *ms)
{
    return guarded([&] {
        if (!d_node_of_part || !d_pids || !d_offsets) throw std::runtime_error("null array");
        Engine &e = engine();
        float t = group_members(d_node_of_part, n, d_node_to_struct, nnodes, nstruct,
                                d_struct_of_part, d_pids, d_offsets, e.stream());
        if (ms) *ms = t;
    });
}

static RamsesArgs ramses_args(const char *repository, const char *nchar, int32_t ncpu,
                              int32_t rver, int32_t family, int32_t stellar, int32_t nthreads)
{
    if (!repository || !nchar) throw std::runtime_error("null path");
    if (family < 0 || family > 2) throw std::runtime_error("family must be 0 (all), 1 (dm) or 2 (star)");
    RamsesArgs A;
    A.repository = repository;
    A.nchar = nchar;
    A.ncpu = ncpu;
    A.rver = rver;
    A.family = family;
    A.stellar = stellar != 0;
    A.nthreads = nthreads;
    return A;
}

int ahop_ramses_count(const char *repository, const char *nchar, int32_t ncpu, int32_t rver,
                      int32_t family, int32_t stellar, int32_t nthreads,
                      ahop_ramses_counts *per_file)
{
    return guarded([&] {
        if (!per_file) throw std::runtime_error("null per_file");
        RamsesArgs A = ramses_args(repository, nchar, ncpu, rver, family, stellar, nthreads);
        std::vector<RamsesFileCount> c(ncpu > 0 ? ncpu : 0);
        ramses_count(A, c.data());
        for (int i = 0; i < ncpu; ++i) per_file[i] = {c[i].ndm, c[i].nstar, c[i].nselected};
    });
}

int ahop_ramses_read(const char *repository, const char *nchar, int32_t ncpu, int32_t rver,
                     int32_t family, int32_t stellar, int32_t nthreads, int64_t ndm_total,
                     int64_t nselected_total, int64_t nstar_total, double scale_l,
                     double scale_t, double *pos, double *vel, double *mass,
                     int64_t *source_pid, double *birth, double *metal, double *m0,
                     double *chem, ahop_ramses_counts *per_file)
{
    return guarded([&] {
        RamsesArgs A = ramses_args(repository, nchar, ncpu, rver, family, stellar, nthreads);
        A.ndm_total = ndm_total;
        A.nselected_total = nselected_total;
        A.nstar_total = nstar_total;
        A.scale_l = scale_l;
        A.scale_t = scale_t;
        A.pos = pos;
        A.vel = vel;
        A.mass = mass;
        A.source_pid = source_pid;
        A.birth = birth;
        A.metal = metal;
        A.m0 = m0;
        A.chem = chem;
        std::vector<RamsesFileCount> c(ncpu > 0 ? ncpu : 0);
        ramses_read(A, c.data());
        if (per_file)
            for (int i = 0; i < ncpu; ++i) per_file[i] = {c[i].ndm, c[i].nstar, c[i].nselected};
    });
}

int ahop_ramses_descriptor(const char *repository, int32_t *present, int32_t *has_m0,
                           int32_t *chem_mask)
{
    return guarded([&] {
        if (!repository) throw std::runtime_error("null path");
        const RamsesDescriptorInfo d = ramses_descriptor(repository);
        if (present) *present = d.present ? 1 : 0;
        if (has_m0) *has_m0 = d.has_m0 ? 1 : 0;
        if (chem_mask) *chem_mask = d.chem_mask;
    });
}

int ahop_halo_stellar_device(int64_t nhalos, const int64_t *offsets, const double *d_pos,
                             const double *d_vel, const double *d_mass, const double *d_density,
                             const uint8_t *d_is_star, const double *d_age,
                             const double *d_metal, const double *d_chem, int32_t nchem,
                             const double *d_props, const ahop_halo_params *params,
                             double massp, double *d_out, double *d_chem_out, float *ms)
{
    return guarded([&] {
        if (!offsets || !d_pos || !d_vel || !d_mass || !d_is_star || !d_age || !d_metal ||
            !d_props || !params || !d_out)
            throw std::runtime_error("null argument");
        static_assert(AHOP_STELLAR_NFIELDS == 33, "stellar field count");
        HaloParamsDev P;
        P.lbox = params->lbox;
        P.gravconst = params->gravconst;
        P.sc = params->sc;
        P.dpmmc = params->dpmmc;
        P.dcell_min = params->dcell_min;
        P.eps_sc = params->eps_sc;
        if (!(P.lbox > 0)) throw std::runtime_error("lbox must be > 0");
        Engine &e = engine();
        float t = halo_stellar_device(offsets, nhalos, d_pos, d_vel, d_mass, d_density, d_is_star,
                                      d_age, d_metal, d_chem, nchem, d_props, P, massp, d_out,
                                      d_chem_out, e.stream());
        if (ms) *ms = t;
    });
}

int ahop_halo_profile_radii_device(int64_t nhalos, const int64_t *offsets, const double *d_pos,
                                   const double *d_mass, const double *d_props, double lbox,
                                   double gravconst, double *d_sdist, double *d_smass,
                                   double *d_scum, double *d_aux, double *d_prof, float *ms)
{
    return guarded([&] {
        if (!offsets || !d_pos || !d_mass || !d_props || !d_sdist || !d_smass || !d_scum ||
            !d_aux || !d_prof)
            throw std::runtime_error("null argument");
        if (!(lbox > 0)) throw std::runtime_error("lbox must be > 0");
        if (nhalos > 0 && offsets[nhalos] > (int64_t)UINT32_MAX)
            throw std::runtime_error("at most 2^32-1 members");
        Engine &e = engine();
        float t = halo_profile_radii(offsets, nhalos, d_pos, d_mass, d_props, lbox, gravconst,
                                     d_sdist, d_smass, d_scum, d_aux, d_prof, e.stream());
        if (ms) *ms = t;
    });
}

int ahop_halo_profile_fits_device(int64_t nhalos, const int64_t *offsets, const double *d_sdist,
                                  const double *d_smass, const double *d_scum,
                                  const double *d_props, const double *d_aux,
                                  const int64_t *edge_offsets, const double *d_edges,
                                  double *d_prof, float *ms)
{
    return guarded([&] {
        if (!offsets || !d_sdist || !d_smass || !d_scum || !d_props || !d_aux || !edge_offsets ||
            !d_prof)
            throw std::runtime_error("null argument");
        if (edge_offsets[nhalos] > 0 && !d_edges) throw std::runtime_error("null edges");
        Engine &e = engine();
        float t = halo_profile_fits(offsets, nhalos, d_sdist, d_smass, d_scum, d_props, d_aux,
                                    edge_offsets, d_edges, d_prof, e.stream());
        if (ms) *ms = t;
    });
}

int ahop_run_device(const double *d_pos, const double *d_mass, int64_t n,
                    const ahop_hot_params *params, double *d_density, int32_t *d_gid,
                    ahop_run_info *info)
{
    return guarded([&] {
        Params p = hot_params(params);
        if (!d_pos || !d_mass) throw std::runtime_error("null input array");
        Engine &e = engine();
        S().dev_n = 0;
        e.run(d_pos, d_mass, n, p, params->saddle != 0);
        S().dev_n = n;
        const DeviceResults &r = e.results();
        hipStream_t st = e.stream();
        if (d_density)
            AHOP_HIP(hipMemcpyAsync(d_density, r.density, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        if (d_gid)
            AHOP_HIP(hipMemcpyAsync(d_gid, r.gid, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
        auto &s = S();
        if (params->saddle) {
            s.dev_sad_offset.resize(r.ngroups + 1);
            s.dev_sad_isad.resize(r.nsaddle);
            s.dev_sad_rho.resize(r.nsaddle);
            AHOP_HIP(hipMemcpyAsync(s.dev_sad_offset.data(), r.sad_offset,
                                    sizeof(int32_t) * (r.ngroups + 1), hipMemcpyDeviceToHost, st));
            if (r.nsaddle > 0) {
                AHOP_HIP(hipMemcpyAsync(s.dev_sad_isad.data(), r.sad_isad,
                                        sizeof(int32_t) * r.nsaddle, hipMemcpyDeviceToHost, st));
                AHOP_HIP(hipMemcpyAsync(s.dev_sad_rho.data(), r.sad_rho, sizeof(double) * r.nsaddle,
                                        hipMemcpyDeviceToHost, st));
            }
        } else {
            s.dev_sad_offset.clear();
            s.dev_sad_isad.clear();
            s.dev_sad_rho.clear();
        }
        AHOP_HIP(hipStreamSynchronize(st));
        fill_info(info, e, params->saddle != 0);
    });
}

int ahop_device_saddle(int32_t *offset, int32_t *isad, double *rho)
{
    return guarded([&] {
        auto &s = S();
        if (offset && !s.dev_sad_offset.empty())
            std::memcpy(offset, s.dev_sad_offset.data(), sizeof(int32_t) * s.dev_sad_offset.size());
        if (isad && !s.dev_sad_isad.empty())
            std::memcpy(isad, s.dev_sad_isad.data(), sizeof(int32_t) * s.dev_sad_isad.size());
        if (rho && !s.dev_sad_rho.empty())
            std::memcpy(rho, s.dev_sad_rho.data(), sizeof(double) * s.dev_sad_rho.size());
    });
}

int ahop_device_hop(int32_t *d_hop, int64_t n)
{
    return guarded([&] {
        if (!d_hop) throw std::runtime_error("null array");
        if (S().dev_n == 0) throw std::runtime_error("no results: call ahop_run_device first");
        if (n != S().dev_n) throw std::runtime_error("n differs from the last ahop_run_device");
        Engine &e = engine();
        AHOP_HIP(hipMemcpyAsync(d_hop, e.results().hop, sizeof(int32_t) * n,
                                hipMemcpyDeviceToDevice, e.stream()));
        AHOP_HIP(hipStreamSynchronize(e.stream()));
    });
}

int ahop_tree_stats(int64_t *out, double *out_ms)
{
    return guarded([&] {
        auto &s = S();
        if (out) std::memcpy(out, s.tree_stats, sizeof(s.tree_stats));
        if (out_ms) std::memcpy(out_ms, s.tree_ms, sizeof(s.tree_ms));
    });
}

int ahop_halo_props_device(int64_t nhalos, const int64_t *offsets, const double *d_pos,
                           const double *d_vel, const double *d_mass, const double *d_density,
                           const ahop_halo_params *params, double *d_out, ahop_halo_info *info)
{
    return guarded([&] {
        if (!params || !offsets || !d_pos || !d_vel || !d_mass || !d_out)
            throw std::runtime_error("null argument");
        static_assert(AHOP_HALO_NFIELDS == HP_NFIELDS, "field count mismatch");
        HaloParamsDev P;
        P.lbox = params->lbox;
        P.hub_pt = params->hub_pt;
        P.gravconst = params->gravconst;
        P.vir_overdens = params->vir_overdens;
        P.rho_mean = params->rho_mean;
        P.cdm = params->cdm;
        P.nshell = params->nshell;
        P.direct_pe_nmax = params->direct_pe_nmax;
        P.sc = params->sc;
        P.dpmmc = params->dpmmc;
        P.dcell_min = params->dcell_min;
        P.eps_sc = params->eps_sc;
        P.light = params->light;
        P.pe_baseline = params->virial_pe_profile;
        P.pe_norm_legacy = params->virial_pe_norm_policy;
        P.shell_half = params->virial_shell_self_mode;
        P.angmom_nshells = params->ang_mom_nshells;
        P.angmom = params->d_ang_mom;
        P.angmom_axes = params->d_ang_axes;
        for (int v : {P.pe_baseline, P.pe_norm_legacy, P.shell_half})
            if (v != 0 && v != 1) throw std::runtime_error("virial switches must be 0 or 1");
        if (!(P.lbox > 0)) throw std::runtime_error("lbox must be > 0");
        Engine &e = engine();
        HaloTuning t = halo_props_device(offsets, nhalos, d_pos, d_vel, d_mass, d_density, P,
                                         d_out, params->tile, params->workgroup, e.stream());
        if (info) {
            info->tile = t.tile;
            info->workgroup = t.bs;
            info->trials = t.trials;
            info->ms = t.ms;
            info->ms_direct_pe = t.sub_ms[0];
            info->ms_virial = t.sub_ms[1];
        }
    });
}

int ahop_create_group_tree(int64_t npart, const double *pos, const double *mass,
                           const double *density, const int32_t *gid, int32_t ngroups,
                           const double *densityg, const int32_t *sad_offset,
                           const int32_t *sad_isad, const double *sad_rho, double rho_threshold,
                           double fudge, double alphap, double epsilon, int32_t nmembthresh,
                           double xlong, int32_t *liste_parts, int32_t *nnodes,
                           double *real_table, int32_t *integer_table)
{
    return guarded([&] {
        if (npart <= 0 || ngroups < 0) throw std::runtime_error("bad sizes");
        TreeInputs ti;
        ti.npart = npart;
        ti.pos = pos;
        ti.mass = mass;
        ti.density = density;
        ti.gid = gid;
        ti.ngroups = ngroups;
        ti.densityg = densityg;
        ti.sad_offset = sad_offset;
        ti.sad_isad = sad_isad;
        ti.sad_rho = sad_rho;
        TreeParams tp;
        tp.rho_threshold = rho_threshold;
        tp.fudge = fudge;
        tp.alphap = alphap;
        tp.epsilon = epsilon;
        tp.nmembthresh = nmembthresh;
        tp.xlong = tp.ylong = tp.zlong = xlong;
        tp.xlongs2 = tp.ylongs2 = tp.zlongs2 = xlong * 0.5;
        tp.optimize_nodes = S().optimize_nodes;
        TreeOutputs to;
        create_group_tree(ti, tp, to);
        keep_tree_stats(S(), to);
        if (liste_parts) std::memcpy(liste_parts, to.liste_parts.data(), sizeof(int32_t) * npart);
        if (nnodes) *nnodes = to.nnodes;
        if (real_table) std::memcpy(real_table, to.real_table.data(), sizeof(double) * to.real_table.size());
        if (integer_table)
            std::memcpy(integer_table, to.integer_table.data(), sizeof(int32_t) * to.integer_table.size());
    });
}

}  // extern "C"
