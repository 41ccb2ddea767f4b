// create_nodes.cpp -- host stage of AdaptaHOP: the structure tree built from the
// device results (group ids, densities, saddle CSR).  It stays on the host, as in
// the reference, and reproduces:
//   create_group_tree             compute_adaptahop.f90:885-983
//   treat_particles(_linkonly)    :2317-2429
//   dev_do_colorize_local         :1188-1242
//   omp_dev_create_nodes_task     :1245-1628
//   reorder_nodes_deterministic   :987-1131
//   node2table                    :550-573
// and, with optimize_nodes, the moment-prefix variant of the node creation:
//   build_opt_prefix_moments      :1639-1745
//   opt_count, treat_particles_opt :1748-1843
//   omp_opt_create_nodes_task     :1845-2278 (its component sums)
//   optimized_needs_exact_fallback :2283-2311
// Child subtrees are OpenMP tasks, as in the reference: the groups, member lists, particles
// and node entries of two sibling subtrees are disjoint, node entries are reserved with an
// atomic counter, and the only scheduling-dependent effect, the pre-reorder node numbering,
// is normalised by reorder_nodes_deterministic.  The components of one level are treated
// in parallel too when there are many (the first level holds every group).
// All sums run in the reference's linked-list order, so the tables match bit for bit.
#include "create_nodes.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <atomic>
#include <vector>

namespace ahop {

namespace {

struct Node {
    int32_t mass = 0, level = 0, mother = 0, firstchild = 0, nsisters = 0, sister = 0;
    double rho_saddle = 0, density = 0, densmax = 0, radius = 0, truemass = 0;
    double position[3] = {0, 0, 0};
};

// one particle of a group list: its index, density, mass and position, so that walking a list
// reads memory in order
struct PartRec {
    double dens, mass, x[3];
    int32_t ipar, pad;
};

// one entry of a group's moment prefix: the k-th densest member's density and the sums over
// the k densest members of mass, mass * position, mass * |position|^2 and density, positions
// unwrapped about the group's densest member
struct Moment {
    double dens, m, s1[3], s2, sd;
};

// indexx f90:220-303 / opt_indexx :306-389 (median-of-three quicksort with an insertion sort
// below 7 elements, whose insertion pass scans down to the first element): ascending order
// of arr[0..n-1] as 1-based positions in indx[1..n].  The order among equal values is this
// algorithm's, so it is restated rather than replaced by std::sort.
void indexx_1based(int32_t n, const double *arr0, std::vector<int32_t> &indx)
{
    constexpr int M = 7, NSTACK = 128;
    auto arr = [arr0](int32_t i) { return arr0[i - 1]; };
    indx.assign((size_t)n + 1, 0);
    int32_t istack[NSTACK + 1];
    for (int32_t j = 1; j <= n; ++j) indx[j] = j;
    int32_t jstack = 0, l = 1, ir = n;
    for (;;) {
        if (ir - l < M) {
            for (int32_t j = l + 1; j <= ir; ++j) {
                const int32_t indxt = indx[j];
                const double a = arr(indxt);
                int32_t i = j - 1;
                for (; i >= 1; --i) {
                    if (arr(indx[i]) <= a) break;
                    indx[i + 1] = indx[i];
                }
                indx[i + 1] = indxt;
            }
            if (jstack == 0) return;
            ir = istack[jstack];
            l = istack[jstack - 1];
            jstack -= 2;
        } else {
            const int32_t k = (l + ir) / 2;
            std::swap(indx[k], indx[l + 1]);
            if (arr(indx[l + 1]) > arr(indx[ir])) std::swap(indx[l + 1], indx[ir]);
            if (arr(indx[l]) > arr(indx[ir])) std::swap(indx[l], indx[ir]);
            if (arr(indx[l + 1]) > arr(indx[l])) std::swap(indx[l + 1], indx[l]);
            int32_t i = l + 1, j = ir;
            const int32_t indxt = indx[l];
            const double a = arr(indxt);
            for (;;) {
                do ++i; while (arr(indx[i]) < a);
                do --j; while (arr(indx[j]) > a);
                if (j < i) break;
                std::swap(indx[i], indx[j]);
            }
            indx[l] = indx[j];
            indx[j] = indxt;
            jstack += 2;
            if (jstack > NSTACK) throw std::runtime_error("indexx: stack too small");
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
}

struct Builder {
    const TreeInputs &in;
    const TreeParams &p;
    int64_t npart;
    int32_t ngroups;
    // each group's particle list (the reference's firstpart / idpart_adapt chains) held as
    // a contiguous segment of packed records in list order, [seg_b[g], seg_e[g]); the
    // record array and liste_parts are process-wide workspaces reused from call to call
    std::vector<PartRec> &rec;
    std::vector<int64_t> seg_b, seg_e;
    std::vector<int32_t> &liste_parts;
    // optimize_nodes: group g's moment prefix is mom[mom_b[g] .. mom_b[g] + mom_n[g])
    std::vector<Moment> &mom;
    std::vector<int64_t> mom_b;
    std::vector<int32_t> mom_n;
    std::vector<int32_t> nhnei, offset1;   // dev_group_nhnei (truncated while colouring), offsets
    std::vector<int32_t> idgroup, igroupid, idgroup_tmp, color;
    std::vector<Node> node;
    int32_t nnodes = 0, nnodesmax = 0;
    int nthreads_ = 1;
    // levels with at least this many colour components treat them in parallel; child
    // subtrees with at least this many groups become tasks
    static constexpr int32_t PAR_COMPONENTS = 16, TASK_MIN_GROUPS = 2;
    // how the parallel branches were taken: levels whose components went to a taskloop,
    // components treated there, and sibling subtrees deferred as tasks
    std::atomic<int64_t> n_taskloop_levels{0}, n_taskloop_components{0}, n_subtree_tasks{0};
    // first error raised inside a task (tasks cannot throw across the parallel region)
    bool failed = false;
    std::string error;

    static void zero_fill(std::vector<int32_t> &v, int64_t n, int nt)
    {
        if ((int64_t)v.size() != n) v.resize(n);
        int32_t *d = v.data();
#pragma omp parallel for num_threads(nt) schedule(static, 65536)
        for (int64_t i = 0; i < n; ++i) d[i] = 0;
    }

    void record_error(const char *what)
    {
#pragma omp critical(ahop_tree_error)
        {
            if (!failed) {
                failed = true;
                error = what;
            }
        }
    }

    Builder(const TreeInputs &i, const TreeParams &pp, std::vector<PartRec> &ws_rec,
            std::vector<int32_t> &ws_liste, std::vector<Moment> &ws_mom)
        : in(i), p(pp), rec(ws_rec), liste_parts(ws_liste), mom(ws_mom)
    {
        npart = in.npart;
        ngroups = in.ngroups;
    }

    double pos(int32_t ipar, int d) const { return in.pos[3 * (int64_t)(ipar - 1) + d]; }
    double dens(int32_t ipar) const { return in.density[ipar - 1]; }
    double densityg(int32_t g) const { return in.densityg[g - 1]; }
    int32_t isad(int32_t idx) const { return in.sad_isad[idx - 1]; }         // 1-based idx
    double rho_sad(int32_t idx) const { return in.sad_rho[idx - 1]; }

    // link_members: the particles above the threshold of each group, in the reference's list
    // order (prepended in index order, so descending index, f90:853-861), packed into that
    // group's segment.  Threads count their index ranges per group, then each particle goes
    // to its slot from the end of its group's segment.
    void link_members(int nt)
    {
        const int T = (int)std::max<int64_t>(1, std::min<int64_t>(nt, npart / 65536));
        std::vector<int64_t> cnt((size_t)T * (ngroups + 1), 0);
        auto active = [&](int64_t ipar) {
            return in.density[ipar - 1] > p.rho_threshold && in.gid[ipar - 1] > 0;
        };
#pragma omp parallel for num_threads(T) schedule(static, 1)
        for (int t = 0; t < T; ++t) {
            const int64_t lo = 1 + npart * t / T, hi = npart * (t + 1) / T;
            int64_t *c = &cnt[(size_t)t * (ngroups + 1)];
            for (int64_t ipar = lo; ipar <= hi; ++ipar)
                if (active(ipar)) ++c[in.gid[ipar - 1]];
        }
        seg_b.assign(ngroups + 2, 0);
        seg_e.assign(ngroups + 2, 0);
        for (int32_t g = 1; g <= ngroups; ++g) {
            int64_t sum = 0;
            for (int t = 0; t < T; ++t) {
                int64_t &c = cnt[(size_t)t * (ngroups + 1) + g];
                const int64_t v = c;
                c = sum;                                    // ascending rank base of thread t
                sum += v;
            }
            seg_b[g] = g > 1 ? seg_e[g - 1] : 0;
            seg_e[g] = seg_b[g] + sum;
        }
        const int64_t nact = ngroups > 0 ? seg_e[ngroups] : 0;
        if ((int64_t)rec.size() < nact) rec.resize(nact);
#pragma omp parallel for num_threads(T) schedule(static, 1)
        for (int t = 0; t < T; ++t) {
            const int64_t lo = 1 + npart * t / T, hi = npart * (t + 1) / T;
            int64_t *rank = &cnt[(size_t)t * (ngroups + 1)];
            for (int64_t ipar = lo; ipar <= hi; ++ipar) {
                if (!active(ipar)) continue;
                const int32_t g = in.gid[ipar - 1];
                PartRec &r = rec[seg_e[g] - 1 - rank[g]++];
                r.dens = in.density[ipar - 1];
                r.mass = in.mass[ipar - 1];
                for (int d = 0; d < 3; ++d) r.x[d] = in.pos[3 * (ipar - 1) + d];
                r.ipar = (int32_t)ipar;
            }
        }
    }

    // treat_particles_linkonly f90:2406-2429: drop the particles at or below rhot, in order
    void treat_linkonly(int32_t igroup, double rhot)
    {
        int64_t w = seg_b[igroup];
        for (int64_t r = seg_b[igroup]; r < seg_e[igroup]; ++r)
            if (rec[r].dens > rhot) rec[w++] = rec[r];
        seg_e[igroup] = w;
    }

    // build_opt_prefix_moments f90:1639-1745, on the lists the first threshold pass left: each
    // group's members in descending density (indexx's ascending order read backwards), the
    // positions unwrapped about the densest member, the running sums in that order
    void build_moments(int nt)
    {
        mom_b.assign(ngroups + 2, 0);
        mom_n.assign(ngroups + 2, 0);
        int64_t total = 0;
        for (int32_t g = 1; g <= ngroups; ++g) {
            mom_b[g] = seg_b[g];
            mom_n[g] = (int32_t)(seg_e[g] - seg_b[g]);
            total = std::max(total, seg_e[g]);
        }
        if ((int64_t)mom.size() < total) mom.resize(total);
        const double L[3] = {p.xlong, p.ylong, p.zlong};
        const double Ls2[3] = {p.xlongs2, p.ylongs2, p.zlongs2};
#pragma omp parallel num_threads(nt)
        {
            std::vector<double> dens;
            std::vector<int32_t> idx;
#pragma omp for schedule(dynamic, 16)
            for (int32_t g = 1; g <= ngroups; ++g) {
                const int32_t n = mom_n[g];
                if (n <= 0 || failed) continue;
                const PartRec *r = rec.data() + seg_b[g];
                dens.resize(n);
                for (int32_t j = 0; j < n; ++j) dens[j] = r[j].dens;
                try {
                    indexx_1based(n, dens.data(), idx);
                } catch (const std::exception &e) {
                    record_error(e.what());
                    continue;
                }
                double a[3];
                for (int d = 0; d < 3; ++d) a[d] = r[idx[n] - 1].x[d];
                double cm = 0.0, cs[3] = {0.0, 0.0, 0.0}, cq = 0.0, cd = 0.0;
                Moment *mg = mom.data() + mom_b[g];
                for (int32_t k = 1; k <= n; ++k) {
                    const PartRec &q = r[idx[n - k + 1] - 1];
                    double u[3];
                    for (int d = 0; d < 3; ++d) {
                        double dx = q.x[d] - a[d];
                        if (dx >= Ls2[d]) dx = dx - L[d];
                        else if (dx < -Ls2[d]) dx = dx + L[d];
                        u[d] = a[d] + dx;
                    }
                    const double xm = q.mass;
                    cm = cm + xm;
                    cs[0] = cs[0] + xm * u[0];
                    cs[1] = cs[1] + xm * u[1];
                    cs[2] = cs[2] + xm * u[2];
                    cq = cq + xm * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                    cd = cd + q.dens;
                    Moment &e = mg[k - 1];
                    e.dens = q.dens;
                    e.m = cm;
                    for (int d = 0; d < 3; ++d) e.s1[d] = cs[d];
                    e.s2 = cq;
                    e.sd = cd;
                }
            }
        }
        if (failed) throw std::runtime_error(error);
    }

    // opt_count + treat_particles_opt f90:1748-1843: the members above rhot are a prefix of the
    // descending order, found by bisection; the sums are the prefix's (the reference's blocks
    // plus residual give the same running sums)
    void treat_moments(int32_t g, double rhot, int32_t &cnt, double &tmass, double s1[3],
                       double &s2, double &sd) const
    {
        const Moment *mg = mom.data() + mom_b[g];
        int32_t lo = 0, hi = mom_n[g];
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) / 2;
            if (mg[mid - 1].dens > rhot) lo = mid;
            else hi = mid - 1;
        }
        cnt = lo;
        if (lo <= 0) {
            tmass = s2 = sd = 0.0;
            s1[0] = s1[1] = s1[2] = 0.0;
            return;
        }
        const Moment &e = mg[lo - 1];
        tmass = e.m;
        for (int d = 0; d < 3; ++d) s1[d] = e.s1[d];
        s2 = e.s2;
        sd = e.sd;
    }

    // optimized_needs_exact_fallback f90:2283-2311: components below 1000 members, or within a
    // relative 1e-8 of one of the selection cuts, are recomputed from the member lists
    bool needs_exact(int32_t mass_loc, double densmoy, double densmaxgroup,
                     double densmoy_comp_max, double rsquare, double rhot) const
    {
        constexpr double rel_margin = 1e-8;
        constexpr int32_t small_component_limit = 1000;
        if (mass_loc <= 0) return false;
        if (mass_loc < small_component_limit) return true;
        const double density_cut = rhot * (1.0 + p.fudge / std::sqrt((double)mass_loc));
        const double alpha_cut = p.alphap * densmoy;
        const double m_density =
            std::fabs(densmoy - density_cut) / std::max(std::fabs(density_cut), 1e-300);
        const double m_comp = std::fabs(densmoy_comp_max - rhot) / std::max(std::fabs(rhot), 1e-300);
        const double m_alpha =
            std::fabs(densmaxgroup - alpha_cut) / std::max(std::fabs(alpha_cut), 1e-300);
        const double m_radius =
            std::fabs(rsquare - p.epsilon) / std::max(std::fabs(p.epsilon), 1e-300);
        return m_density < rel_margin || m_comp < rel_margin || m_alpha < rel_margin ||
               m_radius < rel_margin;
    }

    // treat_particles f90:2317-2401
    void treat(int32_t igroup, double rhot, double posg[3], int32_t &imass, int32_t &igroupref,
               double posref[3], double &rsquare, double &densmoy, double &truemass)
    {
        imass = 0;
        truemass = 0.0;
        rsquare = 0.0;
        densmoy = 0.0;
        posg[0] = posg[1] = posg[2] = 0.0;
        bool first_good = false;
        double densmax = 0, densmin = 0;
        const double L[3] = {p.xlong, p.ylong, p.zlong};
        const double Ls2[3] = {p.xlongs2, p.ylongs2, p.zlongs2};
        int64_t w = seg_b[igroup];
        for (int64_t ri = seg_b[igroup]; ri < seg_e[igroup]; ++ri) {
            const PartRec pr = rec[ri];
            double density_ipar = pr.dens;
            if (density_ipar > rhot) {
                if (!first_good) {
                    if (igroupref == 0) {
                        for (int d = 0; d < 3; ++d) posref[d] = pr.x[d];
                        igroupref = igroup;
                    }
                    first_good = true;
                    densmin = density_ipar;
                    densmax = densmin;
                }
                rec[w++] = pr;                              // the list keeps it, in order
                imass = imass + 1;
                double xmasspart = pr.mass;
                truemass = truemass + xmasspart;
                double pd[3];
                for (int d = 0; d < 3; ++d) {
                    pd[d] = pr.x[d] - posref[d];
                    if (pd[d] >= Ls2[d]) pd[d] = pd[d] - L[d];
                    else if (pd[d] < -Ls2[d]) pd[d] = pd[d] + L[d];
                }
                for (int d = 0; d < 3; ++d) pd[d] = pd[d] + posref[d];
                posg[0] = posg[0] + pd[0] * xmasspart;
                posg[1] = posg[1] + pd[1] * xmasspart;
                posg[2] = posg[2] + pd[2] * xmasspart;
                rsquare = rsquare + xmasspart * (pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2]);
                densmoy = densmoy + density_ipar;
                densmax = std::max(densmax, density_ipar);
                densmin = std::min(densmin, density_ipar);
            }
        }
        seg_e[igroup] = w;
        if (!first_good || densmin <= rhot || densmax != densityg(igroup))
            throw std::runtime_error("treat_particles: group " + std::to_string(igroup) +
                                     " is inconsistent with its density maximum");
    }

    // dev_do_colorize_local f90:1188-1242
    void colorize(int32_t igroup, int32_t igr, double rhot, std::vector<int32_t> &queue,
                  int32_t color_id)
    {
        queue[0] = igroup;
        size_t front = 0, back = 1;
        color[igr] = color_id;
        while (front < back) {
            int32_t g = queue[front++];
            int32_t neig = nhnei[g];
            int32_t base = offset1[g];
            for (int32_t ineig = 1; ineig <= neig; ++ineig) {
                int32_t idx = base + ineig - 1;
                if (rho_sad(idx) > rhot) {
                    int32_t g2 = isad(idx);
                    int32_t igr2 = igroupid[g2];
                    if (color[igr2] == 0) {
                        color[igr2] = color_id;
                        if (back >= queue.size())
                            throw std::runtime_error("colorize: queue overflow");
                        queue[back++] = g2;
                    } else if (color[igr2] != color_id) {
                        throw std::runtime_error(
                            "colorize: saddle connections are not symmetric");
                    }
                } else {
                    nhnei[g] = ineig - 1;
                    break;
                }
            }
        }
    }

    int32_t reserve(int32_t nnew)
    {
        int32_t before;
#pragma omp atomic capture
        {
            before = nnodes;
            nnodes += nnew;
        }
        if (before + nnew > nnodesmax) throw std::runtime_error("create nodes: nnodes > nnodesmax");
        return before + 1;
    }

    // a child subtree as a task body: errors are recorded, not thrown
    void create_nodes_task(double rhot, int32_t inode, int32_t igr1, int32_t igr2)
    {
        if (failed) return;
        try {
            create_nodes(rhot, inode, igr1, igr2, true);
        } catch (const std::exception &e) {
            record_error(e.what());
        }
    }

    // treat_particles over the groups of one colour component, summed as the reference does
    void treat_component(int32_t ic, double rhot, const std::vector<int32_t> &igrposnew,
                         std::vector<double> &posgg, std::vector<int32_t> &massg,
                         std::vector<double> &truemassg, std::vector<double> &densmaxg,
                         std::vector<int32_t> &mass_compg, std::vector<double> &rsquare,
                         std::vector<double> &densmoy, std::vector<char> &ifok)
    {
        if (p.optimize_nodes && treat_component_moments(ic, rhot, igrposnew, posgg, massg,
                                                        truemassg, densmaxg, mass_compg,
                                                        rsquare, densmoy, ifok))
            return;
        double posg[3] = {0, 0, 0};
        double posref[3] = {0, 0, 0};
        int32_t mass_loc = 0;
        double truemass = 0.0, rsquareg = 0.0, densmoyg = 0.0;
        int32_t igr1out = igrposnew[ic - 1] + 1, igr2out = igrposnew[ic];
        double densmaxgroup = -1.0;
        int32_t mass_comp = 0;
        double densmoy_comp_max = -1.0;
        int32_t igroupref = 0;
        for (int32_t igr = igr1out; igr <= igr2out; ++igr) {
            int32_t g = idgroup[igr];
            densmaxgroup = std::max(densmaxgroup, densityg(g));
            double posgtmp[3], rsquaretmp, densmoytmp, truemasstmp;
            int32_t masstmp;
            treat(g, rhot, posgtmp, masstmp, igroupref, posref, rsquaretmp, densmoytmp,
                  truemasstmp);
            posg[0] = posg[0] + posgtmp[0];
            posg[1] = posg[1] + posgtmp[1];
            posg[2] = posg[2] + posgtmp[2];
            rsquareg = rsquareg + rsquaretmp;
            mass_loc = mass_loc + masstmp;
            truemass = truemass + truemasstmp;
            densmoyg = densmoyg + densmoytmp;
            mass_comp = std::max(mass_comp, masstmp);
            if (masstmp > 0) {
                densmoytmp = densmoytmp / (double)masstmp;
                densmoy_comp_max =
                    std::max(densmoy_comp_max,
                             densmoytmp / (1.0 + p.fudge / std::sqrt((double)masstmp)));
            }
        }
        massg[ic] = mass_loc;
        truemassg[ic] = truemass;
        for (int d = 0; d < 3; ++d) posgg[3 * ic + d] = posg[d];
        densmaxg[ic] = densmaxgroup;
        mass_compg[ic] = mass_comp;
        rsquare[ic] = std::sqrt(std::fabs(
            (truemass * rsquareg - (posg[0] * posg[0] + posg[1] * posg[1] + posg[2] * posg[2])) /
            (truemass * truemass)));
        densmoy[ic] = densmoyg / (double)mass_loc;
        ifok[ic] = mass_loc >= p.nmembthresh &&
                   (densmoy[ic] > rhot * (1.0 + p.fudge / std::sqrt((double)mass_loc)) ||
                    densmoy_comp_max > rhot) &&
                   densmaxg[ic] >= p.alphap * densmoy[ic] && rsquare[ic] >= p.epsilon;
    }

    // the component sums of omp_opt_create_nodes_task f90:1980-2036 from the moment prefixes;
    // false (nothing written) when the component needs the exact recomputation
    bool treat_component_moments(int32_t ic, double rhot, const std::vector<int32_t> &igrposnew,
                                 std::vector<double> &posgg, std::vector<int32_t> &massg,
                                 std::vector<double> &truemassg, std::vector<double> &densmaxg,
                                 std::vector<int32_t> &mass_compg, std::vector<double> &rsquare,
                                 std::vector<double> &densmoy, std::vector<char> &ifok)
    {
        double posg[3] = {0, 0, 0};
        int32_t mass_loc = 0, mass_comp = 0;
        double truemass = 0.0, rsquareg = 0.0, densmoyg = 0.0;
        double densmaxgroup = -1.0, densmoy_comp_max = -1.0;
        const int32_t igr1out = igrposnew[ic - 1] + 1, igr2out = igrposnew[ic];
        for (int32_t igr = igr1out; igr <= igr2out; ++igr) {
            const int32_t g = idgroup[igr];
            densmaxgroup = std::max(densmaxgroup, densityg(g));
            int32_t masstmp;
            double truemasstmp, posgtmp[3], rsquaretmp, densmoytmp;
            treat_moments(g, rhot, masstmp, truemasstmp, posgtmp, rsquaretmp, densmoytmp);
            posg[0] = posg[0] + posgtmp[0];
            posg[1] = posg[1] + posgtmp[1];
            posg[2] = posg[2] + posgtmp[2];
            rsquareg = rsquareg + rsquaretmp;
            mass_loc = mass_loc + masstmp;
            truemass = truemass + truemasstmp;
            densmoyg = densmoyg + densmoytmp;
            mass_comp = std::max(mass_comp, masstmp);
            if (masstmp > 0) {
                densmoytmp = densmoytmp / (double)masstmp;
                densmoy_comp_max =
                    std::max(densmoy_comp_max,
                             densmoytmp / (1.0 + p.fudge / std::sqrt((double)masstmp)));
            }
        }
        const double rsq = std::sqrt(std::fabs(
            (truemass * rsquareg - (posg[0] * posg[0] + posg[1] * posg[1] + posg[2] * posg[2])) /
            (truemass * truemass)));
        const double dm = densmoyg / (double)mass_loc;
        if (needs_exact(mass_loc, dm, densmaxgroup, densmoy_comp_max, rsq, rhot)) return false;
        massg[ic] = mass_loc;
        truemassg[ic] = truemass;
        for (int d = 0; d < 3; ++d) posgg[3 * ic + d] = posg[d];
        densmaxg[ic] = densmaxgroup;
        mass_compg[ic] = mass_comp;
        rsquare[ic] = rsq;
        densmoy[ic] = dm;
        ifok[ic] = mass_loc >= p.nmembthresh &&
                   (dm > rhot * (1.0 + p.fudge / std::sqrt((double)mass_loc)) ||
                    densmoy_comp_max > rhot) &&
                   densmaxgroup >= p.alphap * dm && rsq >= p.epsilon;
        // the member lists still drop the particles at or below rhot, so a child paints the
        // members above its parent's threshold (the reference paints the same prefix,
        // opt_count at that threshold)
        for (int32_t igr = igr1out; igr <= igr2out; ++igr) treat_linkonly(idgroup[igr], rhot);
        return true;
    }

    // omp_dev_create_nodes_task f90:1245-1628, run depth first
    void create_nodes(double rhot0, int32_t inode0, int32_t igr1_0, int32_t igr2_0, bool do_paint)
    {
        double rhot = rhot0;
        int32_t inode = inode0, igr1 = igr1_0, igr2 = igr2_0;
        if (do_paint) {
            for (int32_t igr = igr1; igr <= igr2; ++igr) {
                const int32_t g = idgroup[igr];
                for (int64_t r = seg_b[g]; r < seg_e[g]; ++r) liste_parts[rec[r].ipar] = inode;
            }
        }
        const double L[3] = {p.xlong, p.ylong, p.zlong};
        const double Ls2[3] = {p.xlongs2, p.ylongs2, p.zlongs2};
        for (;;) {
            std::vector<int32_t> queue(std::max(igr2 - igr1 + 1, 1));
            for (int32_t igr = igr1; igr <= igr2; ++igr) color[igr] = 0;
            int32_t ncolors = 0;
            for (int32_t igr = igr1; igr <= igr2; ++igr) {
                if (color[igr] == 0) {
                    ++ncolors;
                    colorize(idgroup[igr], igr, rhot, queue, ncolors);
                }
            }
            std::vector<int32_t> igrpos(ncolors + 1, 0), igrinc(ncolors + 1, 0);
            igrpos[0] = igr1 - 1;
            for (int32_t igr = igr1; igr <= igr2; ++igr)
                if (densityg(idgroup[igr]) > rhot) igrpos[color[igr]]++;
            for (int32_t c = 1; c <= ncolors; ++c) igrpos[c] = igrpos[c - 1] + igrpos[c];
            for (int32_t igr = igr1; igr <= igr2; ++igr) {
                int32_t c = color[igr];
                int32_t g = idgroup[igr];
                if (densityg(g) > rhot) {
                    igrinc[c]++;
                    int32_t igr_eff = igrinc[c] + igrpos[c - 1];
                    idgroup_tmp[igr_eff] = g;
                    igroupid[g] = igr_eff;
                }
            }
            igr2 = igrpos[ncolors];
            for (int32_t igr = igr1; igr <= igr2; ++igr) idgroup[igr] = idgroup_tmp[igr];
            int32_t inc_color_tot = 0;
            for (int32_t c = 1; c <= ncolors; ++c)
                if (igrinc[c] > 0) ++inc_color_tot;
            std::vector<int32_t> igrposnew(inc_color_tot + 1);
            igrposnew[0] = igrpos[0];
            inc_color_tot = 0;
            for (int32_t c = 1; c <= ncolors; ++c)
                if (igrinc[c] > 0) igrposnew[++inc_color_tot] = igrpos[c];

            std::vector<double> posgg(3 * (inc_color_tot + 1)), truemassg(inc_color_tot + 1),
                densmaxg(inc_color_tot + 1), rsquare(inc_color_tot + 1),
                densmoy(inc_color_tot + 1);
            std::vector<int32_t> massg(inc_color_tot + 1), mass_compg(inc_color_tot + 1),
                ok_color(inc_color_tot + 1);
            std::vector<char> ifok(inc_color_tot + 1, 0);
            if (inc_color_tot >= PAR_COMPONENTS) {
                n_taskloop_levels += 1;
                n_taskloop_components += inc_color_tot;
                // components touch disjoint groups and particles
#pragma omp taskloop grainsize(4) shared(igrposnew, posgg, massg, truemassg, densmaxg, mass_compg, \
                                         rsquare, densmoy, ifok)
                for (int32_t ic = 1; ic <= inc_color_tot; ++ic) {
                    if (failed) continue;
                    try {
                        treat_component(ic, rhot, igrposnew, posgg, massg, truemassg, densmaxg,
                                        mass_compg, rsquare, densmoy, ifok);
                    } catch (const std::exception &e) {
                        record_error(e.what());
                    }
                }
                if (failed) throw std::runtime_error(error);
            } else {
                for (int32_t ic = 1; ic <= inc_color_tot; ++ic)
                    treat_component(ic, rhot, igrposnew, posgg, massg, truemassg, densmaxg,
                                    mass_compg, rsquare, densmoy, ifok);
            }
            int32_t nsisters = 0, icolor_ref = 0;
            for (int32_t ic = 1; ic <= inc_color_tot; ++ic)
                if (ifok[ic]) {
                    ok_color[++nsisters] = ic;
                    icolor_ref = ic;
                }
            if (nsisters > 1) {
                int32_t inode1 = reserve(nsisters);
                for (int32_t is = 1; is <= nsisters; ++is) {
                    int32_t ic = ok_color[is];
                    int32_t io = inode1 + is - 1;
                    Node &nd = node[io];
                    nd.mother = inode;
                    nd.densmax = densmaxg[ic];
                    nd.sister = is > 1 ? io - 1 : 0;
                    nd.nsisters = nsisters;
                    nd.mass = massg[ic];
                    nd.truemass = truemassg[ic];
                    if (massg[ic] == 0) throw std::runtime_error("create nodes: null mass");
                    double posfin[3];
                    for (int d = 0; d < 3; ++d) posfin[d] = posgg[3 * ic + d] / truemassg[ic];
                    nd.radius = rsquare[ic];
                    nd.density = densmoy[ic];
                    for (int d = 0; d < 3; ++d) {
                        if (posfin[d] >= Ls2[d]) posfin[d] = posfin[d] - L[d];
                        else if (posfin[d] < -Ls2[d]) posfin[d] = posfin[d] + L[d];
                        nd.position[d] = posfin[d];
                    }
                    nd.rho_saddle = rhot;
                    nd.level = node[inode].level + 1;
                }
                node[inode].firstchild = inode1 + nsisters - 1;
                for (int32_t is = 1; is <= nsisters; ++is) {
                    int32_t ic = ok_color[is];
                    const double trho = rhot * (1.0 + p.fudge / std::sqrt((double)mass_compg[ic]));
                    const int32_t child = inode1 + is - 1;
                    const int32_t a = igrposnew[ic - 1] + 1, b = igrposnew[ic];
                    if (b - a + 1 >= TASK_MIN_GROUPS) n_subtree_tasks += 1;
                    // sibling subtrees are independent; small ones run inline
#pragma omp task firstprivate(trho, child, a, b) if (b - a + 1 >= TASK_MIN_GROUPS)
                    create_nodes_task(trho, child, a, b);
                }
                return;
            } else if (nsisters == 1) {
                int32_t ic = icolor_ref;
                double rhotout = rhot * (1.0 + p.fudge / std::sqrt((double)mass_compg[ic]));
                int32_t igr1out = igrposnew[0] + 1, igr2out = igrposnew[inc_color_tot];
                if (igr2out != igr1out) {
                    rhot = rhotout;
                    igr1 = igr1out;
                    igr2 = igr2out;
                    continue;
                }
                node[inode].firstchild = 0;
                return;
            } else {
                node[inode].firstchild = 0;
                return;
            }
        }
    }

    // node_precedes f90:1083-1131
    bool precedes(int32_t a, int32_t b) const
    {
        const Node &x = node[a], &y = node[b];
        if (x.level != y.level) return x.level < y.level;
        if (x.truemass != y.truemass) return x.truemass > y.truemass;
        if (x.mass != y.mass) return x.mass > y.mass;
        if (x.densmax != y.densmax) return x.densmax > y.densmax;
        if (x.density != y.density) return x.density > y.density;
        if (x.rho_saddle != y.rho_saddle) return x.rho_saddle > y.rho_saddle;
        for (int k = 0; k < 3; ++k)
            if (x.position[k] != y.position[k]) return x.position[k] < y.position[k];
        return a < b;
    }

    void reorder()
    {
        if (nnodes <= 1) return;
        std::vector<int32_t> new_to_old(nnodes);
        for (int32_t i = 0; i < nnodes; ++i) new_to_old[i] = i + 1;
        std::sort(new_to_old.begin(), new_to_old.end(),
                  [&](int32_t a, int32_t b) { return precedes(a, b); });
        std::vector<int32_t> old_to_new(nnodes + 1, 0);
        for (int32_t i = 0; i < nnodes; ++i) old_to_new[new_to_old[i]] = i + 1;
        std::vector<Node> sorted(nnodes + 1);
        sorted[0] = node[0];
        for (int32_t i = 1; i <= nnodes; ++i) {
            Node nd = node[new_to_old[i - 1]];
            if (nd.mother > 0) nd.mother = old_to_new[nd.mother];
            if (nd.firstchild > 0) nd.firstchild = old_to_new[nd.firstchild];
            if (nd.sister > 0) nd.sister = old_to_new[nd.sister];
            sorted[i] = nd;
        }
#pragma omp parallel for num_threads(nthreads_) schedule(static, 65536)
        for (int64_t i = 1; i <= npart; ++i)
            if (liste_parts[i] > 0) liste_parts[i] = old_to_new[liste_parts[i]];
        for (int32_t i = 0; i <= nnodes; ++i) node[i] = sorted[i];
    }

    void run(TreeOutputs &out)
    {
        const int nt = p.nthreads > 0 ? p.nthreads : omp_get_max_threads();
        nthreads_ = nt;
        auto now = [] { return omp_get_wtime() * 1e3; };
        double t0 = now();
        link_members(nt);
        nhnei.assign(ngroups + 1, 0);
        offset1.assign(ngroups + 2, 1);
        for (int32_t g = 1; g <= ngroups; ++g) {
            offset1[g] = in.sad_offset[g - 1] + 1;
            nhnei[g] = in.sad_offset[g] - in.sad_offset[g - 1];
        }
        nnodesmax = 2 * ngroups;
        node.assign(nnodesmax + 1, Node{});
        idgroup.assign(ngroups + 1, 0);
        igroupid.assign(ngroups + 1, 0);
        idgroup_tmp.assign(ngroups + 1, 0);
        color.assign(ngroups + 1, 0);
        zero_fill(liste_parts, npart + 1, nt);
        double rhot = p.rho_threshold;
        nnodes = 0;
        node[0].mother = 0;
        double t1 = now();
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64)
        for (int32_t g = 1; g <= ngroups; ++g) treat_linkonly(g, rhot);
        double t2 = now();
        if (p.optimize_nodes) build_moments(nt);   // timed with node creation, as the reference
        node[0].mass = 0;
        node[0].truemass = 0.0;
        node[0].radius = 0.0;
        node[0].density = 0.0;
        double dmx = -HUGE_VAL;
        for (int32_t g = 1; g <= ngroups; ++g) dmx = std::max(dmx, densityg(g));
        node[0].densmax = dmx;
        node[0].rho_saddle = 0.0;
        node[0].level = 0;
        node[0].nsisters = 0;
        node[0].sister = 0;
        for (int32_t g = 1; g <= ngroups; ++g) {
            idgroup[g] = g;
            igroupid[g] = g;
        }
        if (ngroups > 0) {
#pragma omp parallel num_threads(nt)
#pragma omp single
            {
                try {
                    create_nodes(rhot, 0, 1, ngroups, false);
                } catch (const std::exception &e) {
                    record_error(e.what());
                }
            }   // the end of the parallel region waits for every task
            if (failed) throw std::runtime_error(error);
        }
        double t3 = now();
        reorder();
        double t4 = now();
        out.nnodes = nnodes;
        out.liste_parts.resize(npart);            // the caller's recycled buffer: no new pages
        int32_t *dst = out.liste_parts.data();
        const int32_t *src = liste_parts.data() + 1;
#pragma omp parallel for num_threads(nt) schedule(static, 65536)
        for (int64_t i = 0; i < npart; ++i) dst[i] = src[i];
        out.real_table.assign(8 * (size_t)nnodes, 0.0);
        out.integer_table.assign(6 * (size_t)nnodes, 0);
        for (int32_t i = 1; i <= nnodes; ++i) {   // node2table f90:550-573, column-major (8|6, nnodes)
            const Node &nd = node[i];
            double *r = &out.real_table[8 * (size_t)(i - 1)];
            r[0] = nd.rho_saddle;
            r[1] = nd.density;
            r[2] = nd.densmax;
            r[3] = nd.radius;
            r[4] = nd.truemass;
            r[5] = nd.position[0];
            r[6] = nd.position[1];
            r[7] = nd.position[2];
            int32_t *t = &out.integer_table[6 * (size_t)(i - 1)];
            t[0] = nd.mass;
            t[1] = nd.level;
            t[2] = nd.mother;
            t[3] = nd.firstchild;
            t[4] = nd.nsisters;
            t[5] = nd.sister;
        }
        double t5 = now();
        out.ms_link = t1 - t0;
        out.ms_colour = t2 - t1;
        out.ms_nodes = t3 - t2;
        out.ms_reorder = t4 - t3;
        out.ms_output = t5 - t4;
        out.taskloop_levels = n_taskloop_levels;
        out.taskloop_components = n_taskloop_components;
        out.subtree_tasks = n_subtree_tasks;
    }
};

}  // namespace

void create_group_tree(const TreeInputs &in, const TreeParams &p, TreeOutputs &out)
{
    // per-particle workspaces kept for the next call (the module is not re-entrant)
    static std::vector<PartRec> ws_rec;
    static std::vector<int32_t> ws_liste;
    static std::vector<Moment> ws_mom;
    Builder b(in, p, ws_rec, ws_liste, ws_mom);
    b.run(out);
}

}  // namespace ahop
