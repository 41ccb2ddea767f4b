// halo_profiles.hip -- extended profiles of every halo of a CSR member layout:
// compute_extended_profiles_1e (src/compute_halo_props.py:1819-1926) for halos whose members
// are all dark matter.
//
//   radii:  member distances from the halo centre, sorted per halo (a stable radix sort by
//           distance, then one by halo), then per halo the circular-velocity peak over the
//           cumulative mass (vmaxcir, rmaxcir, :1827-1840) and the members inside rvir.
//   fits:   with the log-spaced bin edges of each halo (computed by the caller in NumPy, whose
//           log10 and power decide which end members fall inside, :1862-1863), the weighted
//           histogram, the NFW fit of the binned log-density (cNFW, cNFWerr, :1861-1891) by a
//           bounded Levenberg-Marquardt solve, and the least-squares inner slope (inslope,
//           inslopeerr, :1893-1926).  One wave per halo.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_scan.hpp>

#include "ahop_common.h"
#include "halo_props.h"

namespace ahop {

namespace {

constexpr int PROPS_STRIDE = HP_NFIELDS;   // halo property row (ahop_halo_props_device)
constexpr int P_PX = HP_PX, P_RVIR = HP_RVIR;
constexpr int NPROF = 6;             // vmaxcir, rmaxcir, cNFW, cNFWerr, inslope, inslopeerr
constexpr int NAUX = 4;              // first non-zero slot, members inside rvir, d0, d1
constexpr int MAXBINS = 80;
constexpr double RS_FAIL = 0.01;     // fits ending below this fraction of the innermost bin
                                     // radius count as failed (see k_prof_fit)
constexpr double PI = 3.141592653589793;
constexpr double LN10 = 2.302585092994046;

__device__ __forceinline__ double wrap_strict(double d, double L)
{
    if (d > 0.5 * L) d -= L;
    if (d < -0.5 * L) d += L;
    return d;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// first slot in [lo, hi) with d[slot] > x (upper) or >= x (lower)
__device__ int64_t bound(const double *d, int64_t lo, int64_t hi, double x, bool upper)
{
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (upper ? (d[mid] <= x) : (d[mid] < x))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// distance of every member from its halo's centre (wrapped as correct_for_periodicity_1d),
// with the member's index and halo: the keys of the two sorts below
__global__ void k_prof_keys(const int64_t *off, const double *pos, const double *props,
                            double L, double *dist, int32_t *idx, uint32_t *hid)
{
    const int h = blockIdx.x;
    const double *c = props + (int64_t)h * PROPS_STRIDE + P_PX;
    for (int64_t i = off[h] + threadIdx.x; i < off[h + 1]; i += blockDim.x) {
        const double dx = wrap_strict(pos[3 * i] - c[0], L);
        const double dy = wrap_strict(pos[3 * i + 1] - c[1], L);
        const double dz = wrap_strict(pos[3 * i + 2] - c[2], L);
        dist[i] = sqrt(dx * dx + dy * dy + dz * dz);
        idx[i] = (int32_t)i;
        hid[i] = (uint32_t)h;
    }
}

__global__ void k_prof_gather_hid(const int32_t *idx, const uint32_t *hid, int64_t M,
                                  uint32_t *out)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < M) out[p] = hid[idx[p]];
}

__global__ void k_prof_gather(const int32_t *idx, const double *dist, const double *mass,
                              int64_t M, double *sd, double *sm)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < M) {
        const int32_t i = idx[p];
        sd[p] = dist[i];
        sm[p] = mass[i];
    }
}

// per halo: leading zero distances, the circular-velocity peak over the cumulative mass in
// distance order (first maximum), the members inside rvir and the first / last such distance
// Cumulative mass from a halo's first member at a positive distance, over the members sorted
// by distance: the masses of members at distance 0 masked to 0 and one device-wide segmented
// scan with a head flag at every halo's first member (a workgroup per halo carrying its scan
// chunk by chunk left the largest halo's 271 807 members to one workgroup: 1.75 ms).
__global__ void k_prof_mask(const double *sd, const double *sm, int64_t M, double *mm,
                            uint8_t *head)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= M) return;
    mm[i] = sd[i] > 0.0 ? sm[i] : 0.0;
    head[i] = 0;
}

__global__ void k_prof_heads(const int64_t *off, int64_t nh, uint8_t *head)
{
    const int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (h < nh) head[off[h]] = 1;
}

// Peak circular velocity sqrt(G M(<r) / r) over the members at a positive distance: the first
// maximum of each tile of VMAX_TILE sorted members (k_prof_vmax_tiles), then of each halo's
// tiles in order with the outputs (k_prof_vmax_combine)
constexpr int64_t VMAX_TILE = 8192;
template <int BS>
__global__ __launch_bounds__(BS) void k_prof_vmax_tiles(const int64_t *tile_start,
                                                        const int64_t *tile_end, const double *sd,
                                                        const double *cm, double G,
                                                        double *tbest, int64_t *tbi)
{
    constexpr int NW = BS / 64;
    __shared__ double wbest[NW];
    __shared__ int64_t wbi[NW];
    const int tl = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    double best = -INFINITY;
    int64_t bi = INT64_MAX;
    for (int64_t i = tile_start[tl] + t; i < tile_end[tl]; i += BS) {
        const double d = sd[i];
        if (d > 0.0) {
            const double v = sqrt(G * cm[i] / d);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double b2 = __shfl_xor(best, o);
        const int64_t i2 = __shfl_xor(bi, o);
        if (b2 > best || (b2 == best && i2 < bi)) {
            best = b2;
            bi = i2;
        }
    }
    if (lane == 0) {
        wbest[w] = best;
        wbi[w] = bi;
    }
    __syncthreads();
    if (t == 0) {
        for (int k = 1; k < NW; ++k)
            if (wbest[k] > best || (wbest[k] == best && wbi[k] < bi)) {
                best = wbest[k];
                bi = wbi[k];
            }
        tbest[tl] = best;
        tbi[tl] = bi;
    }
}

__global__ void k_prof_vmax_combine(int64_t nh, const int64_t *off, const int32_t *tile_first,
                                    const double *tbest, const int64_t *tbi, const double *sd,
                                    const double *props, double *aux, double *prof)
{
    const int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (h >= nh) return;
    double best = -INFINITY;
    int64_t bi = INT64_MAX;
    for (int t = tile_first[h]; t < tile_first[h + 1]; ++t)
        if (tbest[t] > best || (tbest[t] == best && tbi[t] < bi)) {
            best = tbest[t];
            bi = tbi[t];
        }
    const int64_t f = off[h], e = off[h + 1];
    const int64_t z = bound(sd, f, e, 0.0, true);      // first distance > 0
    const double rvir = props[h * PROPS_STRIDE + P_RVIR];
    double *o = prof + h * NPROF;
    const bool any = z < e;
    o[0] = any && bi != INT64_MAX ? best : NAN;
    o[1] = any && bi != INT64_MAX ? sd[bi] : NAN;
    for (int q = 2; q < NPROF; ++q) o[q] = NAN;
    const int64_t nin = rvir > 0 ? bound(sd, z, e, rvir, true) - z : 0;
    double *a = aux + h * NAUX;
    a[0] = (double)(z - f);
    a[1] = (double)nin;
    a[2] = nin > 0 ? sd[z] : NAN;
    a[3] = nin > 0 ? sd[z + nin - 1] : NAN;
}

// the NFW model residual (log_nfw - y) * w and its derivative in rs (:1864-1866)
__device__ __forceinline__ void nfw_point(double r, double y, double w, double rs, double a,
                                          double &res, double &drs)
{
    const double x = r / rs;
    res = (a - log10(x) - 2.0 * log10(1.0 + x) - y) * w;
    drs = (1.0 + 3.0 * x) / ((1.0 + x) * rs * LN10) * w;
}

// one wave per halo
__global__ __launch_bounds__(64) void k_prof_fit(const int64_t *off, const double *sd,
                                                 const double *sm, const double *cm,
                                                 const double *props,
                                                 const double *aux, const int64_t *eoff,
                                                 const double *edges, double *prof)
{
    __shared__ int64_t cum[MAXBINS + 1];
    __shared__ double bmass[MAXBINS];
    __shared__ double R[MAXBINS], Y[MAXBINS], W[MAXBINS];
    const int h = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t ne = eoff[h + 1] - eoff[h];
    if (ne < 2) return;                                 // not eligible for the fits
    const int nb = (int)ne - 1;
    const double *b = edges + eoff[h];
    const double *a4 = aux + (int64_t)h * NAUX;
    const int64_t lo = off[h] + (int64_t)a4[0];
    const int64_t hi = lo + (int64_t)a4[1];             // members inside rvir: [lo, hi)
    const double rvir = props[(int64_t)h * PROPS_STRIDE + P_RVIR];
    double *o = prof + (int64_t)h * NPROF;

    // np.histogram with explicit edges: counts below each edge (<= for the last edge); the
    // bin masses as differences of the cumulative mass, as np.histogram computes weighted bins
    for (int j = lane; j <= nb; j += 64) cum[j] = bound(sd, lo, hi, b[j], j == nb);
    __syncthreads();
    for (int j = lane; j < nb; j += 64) {
        const double c1 = cum[j + 1] > lo ? cm[cum[j + 1] - 1] : 0.0;
        const double c0 = cum[j] > lo ? cm[cum[j] - 1] : 0.0;
        bmass[j] = c1 - c0;
    }
    __syncthreads();
    // valid bins, compacted in order
    int nv = 0;
    for (int base = 0; base < nb; base += 64) {
        const int j = base + lane;
        double vol = 0.0, m = 0.0;
        int64_t cnt = 0;
        if (j < nb) {
            const double b0 = b[j], b1 = b[j + 1];
            vol = 4.0 / 3.0 * PI * (b1 * b1 * b1 - b0 * b0 * b0);
            m = bmass[j];
            cnt = cum[j + 1] - cum[j];
        }
        const bool valid = j < nb && cnt > 0 && m > 0.0 && vol > 0.0;
        const uint64_t mask = __ballot(valid);
        if (valid) {
            const int k = nv + __popcll(mask & ((1ull << lane) - 1));
            R[k] = 0.5 * (b[j + 1] + b[j]);
            Y[k] = log10(m / vol);
            W[k] = sqrt((double)cnt);                   // 1 / sigma, sigma = 1/sqrt(count)
        }
        nv += __popcll(mask);
    }
    __syncthreads();
    double cnfw = NAN;
    if (nv >= 4) {
        // p0 = [max(rvir/10, radius[0]), log10(rho[0])], bounds rs in [0, rvir]
        double rs = fmax(rvir / 10.0, R[0]), a = Y[0];
        const double rs_min = rvir * 1e-12;
        auto cost_at = [&](double prs, double pa) {
            double c = 0.0;
            for (int k = lane; k < nv; k += 64) {
                double r, d;
                nfw_point(R[k], Y[k], W[k], prs, pa, r, d);
                c += r * r;
            }
            return wave_sum(c);
        };
        double cost = cost_at(rs, a);
        double lam = 1e-3;
        for (int it = 0; it < 500; ++it) {
            double A00 = 0, A01 = 0, A11 = 0, g0 = 0, g1 = 0;
            for (int k = lane; k < nv; k += 64) {
                double r, d;
                nfw_point(R[k], Y[k], W[k], rs, a, r, d);
                A00 += d * d;
                A01 += d * W[k];
                A11 += W[k] * W[k];
                g0 += d * r;
                g1 += W[k] * r;
            }
            A00 = wave_sum(A00);
            A01 = wave_sum(A01);
            A11 = wave_sum(A11);
            g0 = wave_sum(g0);
            g1 = wave_sum(g1);
            bool accepted = false, converged = false;
            while (lam < 1e30) {
                const double m00 = A00 * (1.0 + lam), m11 = A11 * (1.0 + lam);
                const double det = m00 * m11 - A01 * A01;
                if (!(det > 0)) {
                    lam *= 10.0;
                    continue;
                }
                const double d0 = -(m11 * g0 - A01 * g1) / det;
                const double d1 = -(m00 * g1 - A01 * g0) / det;
                const double rs2 = fmin(fmax(rs + d0, rs_min), rvir), a2 = a + d1;
                // a step below rounding level: converged (evaluating it would only be rejected
                // by rounding, raising lam step after step)
                if (fabs(rs2 - rs) <= 1e-14 * rs && fabs(d1) <= 1e-14 * fmax(1.0, fabs(a))) {
                    converged = true;
                    accepted = true;
                    break;
                }
                const double c2 = cost_at(rs2, a2);
                if (c2 <= cost) {
                    converged = fabs(rs2 - rs) <= 1e-15 * rs && fabs(a2 - a) <= 1e-15 * fmax(1.0, fabs(a));
                    rs = rs2;
                    a = a2;
                    cost = c2;
                    lam = fmax(lam * 0.1, 1e-15);
                    accepted = true;
                    break;
                }
                lam *= 10.0;
            }
            if (!accepted || converged) break;
        }
        // covariance: inv(J^T J) scaled by the residual variance (absolute_sigma=False)
        double A00 = 0, A01 = 0, A11 = 0;
        for (int k = lane; k < nv; k += 64) {
            double r, d;
            nfw_point(R[k], Y[k], W[k], rs, a, r, d);
            A00 += d * d;
            A01 += d * W[k];
            A11 += W[k] * W[k];
        }
        A00 = wave_sum(A00);
        A01 = wave_sum(A01);
        A11 = wave_sum(A11);
        const double det = A00 * A11 - A01 * A01;
        const double var = det > 0 ? A11 / det * cost / (nv - 2) : INFINITY;
        const double rs_err = sqrt(var);
        // An optimum far inside the innermost bin makes the model a pure r^-3 power law over
        // every bin: scipy's bounded TRF then creeps toward rs = 0 until its 200 evaluations
        // run out and curve_fit raises, which the reference turns into NaN (:1888-1890).
        if (rs >= RS_FAIL * R[0]) {
            cnfw = rvir / rs;
            if (lane == 0) {
                o[2] = cnfw;
                o[3] = rs > 0 ? cnfw * sqrt((rs_err / rs) * (rs_err / rs)) : NAN;
            }
        }
    }
    // inner slope inside min(rvir / cNFW, rvir) (:1893-1926): per-member shell densities
    const double rs_eff = (isfinite(cnfw) && cnfw > 0) ? rvir / cnfw : rvir;
    const double rcut = fmin(rs_eff, rvir);
    const int64_t kend = bound(sd, lo, hi, rcut, false);          // members with d < rcut
    if (kend - lo < 4) return;
    auto point = [&](int64_t i, double &x, double &y) {
        const double r = sd[i], m = sm[i];
        const double v = 4.0 / 3.0 * PI * (r * r * r);
        const double rp = i > lo ? sd[i - 1] : 0.0;
        const double vp = i > lo ? 4.0 / 3.0 * PI * (rp * rp * rp) : 0.0;
        const double dv = v - vp;
        const double rel = i > lo ? (r - rp) / fmax(r, 2.2250738585072014e-308) : 1.0;
        x = log10(r / rvir);
        y = log10(m / dv);
        return dv > 0.0 && m > 0.0 && rel > 1.0e-9;
    };
    double n = 0, sx = 0, sy = 0;
    for (int64_t i = lo + lane; i < kend; i += 64) {
        double x, y;
        if (point(i, x, y)) {
            n += 1.0;
            sx += x;
            sy += y;
        }
    }
    n = wave_sum(n);
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (n < 4) return;
    const double mx = sx / n, my = sy / n;
    double sxx = 0, sxy = 0;
    for (int64_t i = lo + lane; i < kend; i += 64) {
        double x, y;
        if (point(i, x, y)) {
            sxx += (x - mx) * (x - mx);
            sxy += (x - mx) * (y - my);
        }
    }
    sxx = wave_sum(sxx);
    sxy = wave_sum(sxy);
    const double slope = sxy / sxx, icpt = my - slope * mx;
    double rss = 0;
    for (int64_t i = lo + lane; i < kend; i += 64) {
        double x, y;
        if (point(i, x, y)) {
            const double r = y - (slope * x + icpt);
            rss += r * r;
        }
    }
    rss = wave_sum(rss);
    if (lane == 0) {
        o[4] = slope;
        o[5] = sqrt(rss / (n - 2) / sxx);
    }
}

}  // namespace

float halo_profile_radii(const int64_t *h_off, int64_t nh, const double *d_pos,
                         const double *d_mass, const double *d_props, double L, double G,
                         double *d_sdist, double *d_smass, double *d_scum, double *d_aux,
                         double *d_prof, hipStream_t s)
{
    if (nh < 1 || nh > INT32_MAX / 2) throw std::runtime_error("need 1 <= nhalos < 2^30");
    if (h_off[0] != 0) throw std::runtime_error("offsets must start at 0");
    for (int64_t h = 0; h < nh; ++h)
        if (h_off[h + 1] <= h_off[h]) throw std::runtime_error("every halo needs at least one member");
    const int64_t M = h_off[nh];
    if (M > INT32_MAX) throw std::runtime_error("at most 2^31-1 members");
    // sorted per halo by two stable device-wide radix sorts: by distance, then by halo (the
    // second keeps the first's order within a halo), so no halo waits on one workgroup
    static DevBuf doff, ddist, dd2, didx, di2, dhid, dh2, temp;
    // tiles of the peak-velocity search (the host lists live until the synchronisation below)
    std::vector<int32_t> vtf(nh + 1);
    std::vector<int64_t> vts, vte;
    int64_t *d_off = (int64_t *)doff.ensure(sizeof(int64_t) * (nh + 1));
    double *dist = (double *)ddist.ensure(sizeof(double) * M);
    double *dist_s = (double *)dd2.ensure(sizeof(double) * M);
    int32_t *idx = (int32_t *)didx.ensure(sizeof(int32_t) * M);
    int32_t *idx_s = (int32_t *)di2.ensure(sizeof(int32_t) * M);
    uint32_t *hid = (uint32_t *)dhid.ensure(sizeof(uint32_t) * M);
    uint32_t *hid_s = (uint32_t *)dh2.ensure(sizeof(uint32_t) * M);
    int hbits = 1;
    while (hbits < 32 && ((nh - 1) >> hbits) != 0) ++hbits;
    hipEvent_t e0, e1;
    AHOP_HIP(hipEventCreate(&e0));
    AHOP_HIP(hipEventCreate(&e1));
    AHOP_HIP(hipMemcpyAsync(d_off, h_off, sizeof(int64_t) * (nh + 1), hipMemcpyHostToDevice, s));
    AHOP_HIP(hipEventRecord(e0, s));
    const unsigned gb = (unsigned)((M + 255) / 256);
    hipLaunchKernelGGL(k_prof_keys, dim3((unsigned)nh), dim3(256), 0, s, d_off, d_pos, d_props,
                       L, dist, idx, hid);
    size_t tb1 = 0, tb2 = 0;
    AHOP_HIP(rocprim::radix_sort_pairs(nullptr, tb1, dist, dist_s, idx, idx_s, (size_t)M, 0, 64, s));
    AHOP_HIP(rocprim::radix_sort_pairs(nullptr, tb2, hid_s, hid, idx_s, idx, (size_t)M, 0, hbits, s));
    void *tmp = temp.ensure(std::max(tb1, tb2));
    AHOP_HIP(rocprim::radix_sort_pairs(tmp, tb1, dist, dist_s, idx, idx_s, (size_t)M, 0, 64, s));
    hipLaunchKernelGGL(k_prof_gather_hid, dim3(gb), dim3(256), 0, s, idx_s, hid, (int64_t)M, hid_s);
    AHOP_HIP(rocprim::radix_sort_pairs(tmp, tb2, hid_s, hid, idx_s, idx, (size_t)M, 0, hbits, s));
    hipLaunchKernelGGL(k_prof_gather, dim3(gb), dim3(256), 0, s, idx, dist, d_mass, (int64_t)M,
                       d_sdist, d_smass);
    {
        static DevBuf bmm, bhead, bscan, btf, bts, bte, bbest, bbi;
        double *mm = (double *)bmm.ensure(sizeof(double) * M);
        uint8_t *head = (uint8_t *)bhead.ensure(M);
        hipLaunchKernelGGL(k_prof_mask, dim3(gb), dim3(256), 0, s, d_sdist, d_smass, (int64_t)M,
                           mm, head);
        hipLaunchKernelGGL(k_prof_heads, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s,
                           d_off, nh, head);
        size_t tb = 0;
        AHOP_HIP(rocprim::segmented_inclusive_scan(nullptr, tb, mm, d_scum, head, (size_t)M,
                                                   rocprim::plus<double>(), s));
        AHOP_HIP(rocprim::segmented_inclusive_scan(bscan.ensure(tb), tb, mm, d_scum, head,
                                                   (size_t)M, rocprim::plus<double>(), s));
        for (int64_t h = 0; h < nh; ++h) {
            vtf[h] = (int32_t)vts.size();
            for (int64_t a = h_off[h]; a < h_off[h + 1]; a += VMAX_TILE) {
                vts.push_back(a);
                vte.push_back(std::min(a + VMAX_TILE, h_off[h + 1]));
            }
        }
        vtf[nh] = (int32_t)vts.size();
        const int64_t nt = (int64_t)vts.size();
        int32_t *d_tf = (int32_t *)btf.ensure(sizeof(int32_t) * (nh + 1));
        int64_t *d_ts = (int64_t *)bts.ensure(sizeof(int64_t) * nt);
        int64_t *d_te = (int64_t *)bte.ensure(sizeof(int64_t) * nt);
        double *d_best = (double *)bbest.ensure(sizeof(double) * nt);
        int64_t *d_bi = (int64_t *)bbi.ensure(sizeof(int64_t) * nt);
        AHOP_HIP(hipMemcpyAsync(d_tf, vtf.data(), sizeof(int32_t) * (nh + 1), hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_ts, vts.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_te, vte.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_prof_vmax_tiles<256>, dim3((unsigned)nt), dim3(256), 0, s, d_ts, d_te,
                           d_sdist, d_scum, G, d_best, d_bi);
        hipLaunchKernelGGL(k_prof_vmax_combine, dim3((unsigned)((nh + 127) / 128)), dim3(128), 0, s,
                           nh, d_off, d_tf, d_best, d_bi, d_sdist, d_props, d_aux, d_prof);
    }
    AHOP_HIP(hipEventRecord(e1, s));
    AHOP_HIP(hipStreamSynchronize(s));
    float ms = 0;
    AHOP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

float halo_profile_fits(const int64_t *h_off, int64_t nh, const double *d_sdist,
                        const double *d_smass, const double *d_scum, const double *d_props,
                        const double *d_aux,
                        const int64_t *h_eoff, const double *d_edges, double *d_prof,
                        hipStream_t s)
{
    if (nh < 1 || nh > INT32_MAX / 2) throw std::runtime_error("need 1 <= nhalos < 2^30");
    for (int64_t h = 0; h < nh; ++h) {
        const int64_t ne = h_eoff[h + 1] - h_eoff[h];
        if (ne != 0 && (ne < 9 || ne > MAXBINS + 1))
            throw std::runtime_error("a halo's bin edges must number 0 or 9 .. 81");
    }
    static DevBuf doff, deo;
    int64_t *d_off = (int64_t *)doff.ensure(sizeof(int64_t) * (nh + 1));
    int64_t *d_eoff = (int64_t *)deo.ensure(sizeof(int64_t) * (nh + 1));
    hipEvent_t e0, e1;
    AHOP_HIP(hipEventCreate(&e0));
    AHOP_HIP(hipEventCreate(&e1));
    AHOP_HIP(hipMemcpyAsync(d_off, h_off, sizeof(int64_t) * (nh + 1), hipMemcpyHostToDevice, s));
    AHOP_HIP(hipMemcpyAsync(d_eoff, h_eoff, sizeof(int64_t) * (nh + 1), hipMemcpyHostToDevice, s));
    AHOP_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_prof_fit, dim3((unsigned)nh), dim3(64), 0, s, d_off, d_sdist, d_smass,
                       d_scum, d_props, d_aux, d_eoff, d_edges, d_prof);
    AHOP_HIP(hipEventRecord(e1, s));
    AHOP_HIP(hipStreamSynchronize(s));
    float ms = 0;
    AHOP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

}  // namespace ahop
