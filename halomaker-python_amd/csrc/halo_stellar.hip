// halo_stellar.hip -- per-halo stellar properties (the stellar payload of
// src/compute_halo_props.py, mode 'full') for every halo of a CSR member layout:
//
//   det_mass_17 :765-780 and det_center_18 :963-966: member counts and masses by type, the
//     contamination mass, the dispersions of dark matter and stars about the halo's mean
//     velocity (k_type_sums, one pass over all members);
//   refine_centers_1b :1569-1593, compute_ang_mom_19 :806-822, r_halos_1a :842-849: the
//     stellar centre, L* and r*, by the batched halo properties (light) over each halo's stars,
//     gathered in member order into a star CSR (k_star_flags, a scan, k_star_scatter);
//   compute_stellar_1d :1635-1751: r50 / r90 from the cumulative mass in distance order (a
//     sequential sum per halo, as np.cumsum, after two stable radix sorts), mass-weighted age
//     and metallicity, the star formation rates, and vrot / sig3d / sigcyl over all stars,
//     inside r50 and inside r90 (_stellar_kinematics_1d :1604-1633) (k_stellar), and the
//     mass-weighted chemistry per element (k_stellar_chem).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "ahop_common.h"
#include "halo_props.h"

namespace ahop {

namespace {

constexpr int BS = 256, NW = BS / 64;
constexpr int PROPS_STRIDE = HP_NFIELDS;
enum StellarField {
    S_NDM = 0, S_NSTAR, S_MDM, S_MSTAR, S_MCONTAM, S_SIGDM, S_SIGST, S_PX, S_PY, S_PZ,
    S_LX, S_LY, S_LZ, S_RST, S_R50, S_R90, S_AGE, S_METAL, S_SFR, S_SFR50, S_SFR90, S_SFR10,
    S_SFR10_50, S_SFR10_90, S_VROT, S_SIG3D, S_SIGCYL, S_NFIELDS = S_VROT + 9
};
constexpr double DBL_EPS = 2.220446049250313e-16;

__device__ __forceinline__ double wrap_strict(double d, double L)
{
    if (d > 0.5 * L) d -= L;
    if (d < -0.5 * L) d += L;
    return d;
}

// block-wide sums: wave shuffles, then the waves in order through LDS (sh: NQ * NW doubles)
template <int NQ>
__device__ void block_sums(double (&v)[NQ], double *sh)
{
    for (int q = 0; q < NQ; ++q)
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < NQ; ++q) sh[q * NW + (threadIdx.x >> 6)] = v[q];
    __syncthreads();
    for (int q = 0; q < NQ; ++q) {
        double a = sh[q * NW];
        for (int w = 1; w < NW; ++w) a += sh[q * NW + w];
        v[q] = a;
    }
    __syncthreads();
}

// counts and masses by type, contamination, dispersions about the halo's mean velocity;
// the row's other fields get the values clear_halo leaves without stars
// Per-tile sums of a halo's members by type (a halo's members cut into tiles of at most
// TYPE_TILE, so that the largest halo is not one workgroup's walk), combined per halo in tile
// order by k_type_combine.
constexpr int64_t TYPE_TILE = 8192;
__global__ __launch_bounds__(BS) void k_type_sums(const int32_t *tile_halo,
                                                  const int64_t *tile_start,
                                                  const int64_t *tile_end, const double *vel,
                                                  const double *mass, const uint8_t *is_star,
                                                  const double *props, double massp, double *part)
{
    __shared__ double sh[7 * NW];
    const int t = blockIdx.x;
    const int h = tile_halo[t];
    const double *vb = props + (int64_t)h * PROPS_STRIDE + HP_VX;
    const double vx = vb[0], vy = vb[1], vz = vb[2];
    double v[7] = {0, 0, 0, 0, 0, 0, 0};   // ndm, nstar, mdm, m*, mcontam, dv2 dm, dv2 star
    for (int64_t i = tile_start[t] + threadIdx.x; i < tile_end[t]; i += BS) {
        const double m = mass[i];
        const double dvx = vel[3 * i] - vx, dvy = vel[3 * i + 1] - vy, dvz = vel[3 * i + 2] - vz;
        const double dv2 = m * (dvx * dvx + dvy * dvy + dvz * dvz);
        if (is_star[i]) {
            v[1] += 1.0;
            v[3] += m;
            v[6] += dv2;
        } else {
            v[0] += 1.0;
            v[2] += m;
            if (m > massp * 1.00001) v[4] += m;
            v[5] += dv2;
        }
    }
    block_sums<7>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 7; ++q) part[(int64_t)t * 7 + q] = v[q];
}

__global__ void k_type_combine(int64_t nh, const int32_t *tile_first, const double *part,
                               int64_t *scount, double *out)
{
    const int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (h >= nh) return;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int t = tile_first[h]; t < tile_first[h + 1]; ++t)
        for (int q = 0; q < 7; ++q) v[q] += part[(int64_t)t * 7 + q];
    double *o = out + h * S_NFIELDS;
    for (int q = 0; q < S_NFIELDS; ++q) o[q] = NAN;
    for (int q = S_SFR; q <= S_SFR10_90; ++q) o[q] = 0.0;
    o[S_NDM] = v[0];
    o[S_NSTAR] = v[1];
    o[S_MDM] = v[2];
    o[S_MSTAR] = v[3];
    o[S_MCONTAM] = v[4];
    o[S_SIGDM] = v[2] > 0 ? sqrt(v[5] / v[2]) : 0.0;
    o[S_SIGST] = v[3] > 0 ? sqrt(v[6] / v[3]) : 0.0;
    scount[h] = (int64_t)v[1];
}

// Each halo's stars, in member order, go into the star arrays from soff[h].  The halos' member
// ranges are consecutive, so a star's place is the number of stars before it among all members:
// an exclusive scan of the star flags (a workgroup per halo walking its members took 1.5 ms at
// the 50 000-halo workload, the largest halo's 271 807 members on one workgroup).
__global__ void k_star_flags(const uint8_t *is_star, int64_t M, int32_t *flag)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < M) flag[i] = is_star[i] ? 1 : 0;
}

__global__ void k_star_scatter(const uint8_t *is_star, const int32_t *place, int64_t M,
                               const double *pos, const double *vel, const double *mass,
                               const double *dens, const double *age, const double *metal,
                               const double *chem, int nchem, double *spos, double *svel,
                               double *smass, double *sdens, double *sage, double *smetal,
                               double *schem)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= M || !is_star[i]) return;
    const int64_t j = place[i];
    for (int d = 0; d < 3; ++d) {
        spos[3 * j + d] = pos[3 * i + d];
        svel[3 * j + d] = vel[3 * i + d];
    }
    smass[j] = mass[i];
    sdens[j] = dens[i];
    sage[j] = age[i];
    smetal[j] = metal[i];
    for (int c = 0; c < nchem; ++c) schem[j * nchem + c] = chem[i * nchem + c];
}

// squared distance of every star from its stellar centre, with its index and star halo
__global__ void k_star_keys(const int64_t *soff, const double *spos, const double *sprops,
                            double L, double *dr2, int32_t *idx, uint32_t *hid)
{
    const int k = blockIdx.x;
    const double *c = sprops + (int64_t)k * PROPS_STRIDE + HP_PX;
    for (int64_t i = soff[k] + threadIdx.x; i < soff[k + 1]; i += blockDim.x) {
        const double dx = wrap_strict(spos[3 * i] - c[0], L);
        const double dy = wrap_strict(spos[3 * i + 1] - c[1], L);
        const double dz = wrap_strict(spos[3 * i + 2] - c[2], L);
        dr2[i] = dx * dx + dy * dy + dz * dz;
        idx[i] = (int32_t)i;
        hid[i] = (uint32_t)k;
    }
}

__global__ void k_gather_u32(const int32_t *idx, const uint32_t *src, int64_t n, uint32_t *dst)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[idx[p]];
}

__global__ void k_gather_sorted(const int32_t *idx, const double *dr2, const double *smass,
                                int64_t n, double *sd2, double *sm)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < n) {
        sd2[p] = dr2[idx[p]];
        sm[p] = smass[idx[p]];
    }
}

// r50 / r90 (:1662-1668): the cumulative mass in distance order summed left to right as
// np.cumsum does (one thread over chunks staged in LDS by the block), then the first slot
// reaching half (90%) of the total by binary search, as np.searchsorted
constexpr int CUM_CHUNK = 2048;

__global__ __launch_bounds__(BS) void k_star_r50(const int64_t *soff, const double *sd2,
                                                 const double *sm, double *cum, double *r50,
                                                 double *r90)
{
    __shared__ double buf[CUM_CHUNK];
    __shared__ double carry;
    const int k = blockIdx.x;
    const int64_t f = soff[k], e = soff[k + 1];
    if (threadIdx.x == 0) carry = 0.0;
    for (int64_t base = f; base < e; base += CUM_CHUNK) {
        const int len = (int)((e - base) < CUM_CHUNK ? (e - base) : CUM_CHUNK);
        for (int j = threadIdx.x; j < len; j += BS) buf[j] = sm[base + j];
        __syncthreads();
        if (threadIdx.x == 0) {
            double c = carry;
#pragma unroll 8
            for (int j = 0; j < len; ++j) {
                c += buf[j];
                buf[j] = c;
            }
            carry = c;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < len; j += BS) cum[base + j] = buf[j];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double total = carry;
        auto first_at_least = [&](double v) {
            int64_t lo = f, hi = e;
            while (lo < hi) {
                const int64_t mid = lo + (hi - lo) / 2;
                if (cum[mid] < v)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            return lo < e ? lo : e - 1;
        };
        r50[k] = sqrt(sd2[first_at_least(0.5 * total)]);
        r90[k] = sqrt(sd2[first_at_least(0.9 * total)]);
    }
}

// stellar chemistry (:1671-1679): per element the mass-weighted mean over the stars with a
// finite positive mass and a finite value (NaN when there are none)
constexpr int MAX_CHEM = 9;

__global__ __launch_bounds__(BS) void k_stellar_chem(const int32_t *hlist, const int64_t *soff,
                                                     const double *smass, const double *schem,
                                                     int nchem, double *chem_out)
{
    __shared__ double sh[2 * MAX_CHEM * NW];
    const int k = blockIdx.x;
    double v[2 * MAX_CHEM] = {0};
    for (int64_t i = soff[k] + threadIdx.x; i < soff[k + 1]; i += BS) {
        const double m = smass[i];
        if (!(isfinite(m) && m > 0)) continue;
        for (int c = 0; c < nchem; ++c) {
            const double x = schem[i * nchem + c];
            if (isfinite(x)) {
                v[2 * c] += x * m;
                v[2 * c + 1] += m;
            }
        }
    }
    block_sums<2 * MAX_CHEM>(v, sh);
    if (threadIdx.x == 0)
        for (int c = 0; c < nchem; ++c)
            chem_out[(int64_t)hlist[k] * nchem + c] = v[2 * c + 1] > 0 ? v[2 * c] / v[2 * c + 1] : NAN;
}

// compute_stellar_1d's sums for one star halo per block
__global__ __launch_bounds__(BS) void k_stellar(const int32_t *hlist, const int64_t *soff,
                                                const double *spos, const double *svel,
                                                const double *smass, const double *sage,
                                                const double *smetal, const double *sprops,
                                                const double *r50v, const double *r90v, double L,
                                                double *out)
{
    __shared__ double sh[24 * NW];
    const int k = blockIdx.x;
    const int64_t f = soff[k], e = soff[k + 1];
    const int64_t n = e - f;
    const double *sp = sprops + (int64_t)k * PROPS_STRIDE;
    const double cx = sp[HP_PX], cy = sp[HP_PY], cz = sp[HP_PZ];
    const double Lx = sp[HP_LX], Ly = sp[HP_LY], Lz = sp[HP_LZ];
    const double r50 = r50v[k], r90 = r90v[k];
    const double r50s = r50 * r50, r90s = r90 * r90;
    auto rel = [&](int64_t i, double &dx, double &dy, double &dz) {
        dx = wrap_strict(spos[3 * i] - cx, L);
        dy = wrap_strict(spos[3 * i + 1] - cy, L);
        dz = wrap_strict(spos[3 * i + 2] - cz, L);
        return dx * dx + dy * dy + dz * dz;
    };
    // pass 1: mass, age, metal, star formation, momentum
    double a[12] = {0};
    for (int64_t i = f + threadIdx.x; i < e; i += BS) {
        double dx, dy, dz;
        const double d2 = rel(i, dx, dy, dz);
        const double m = smass[i], ag = sage[i];
        const bool in50 = d2 < r50s, in90 = d2 < r90s;
        a[0] += m;
        a[1] += ag * m;
        a[2] += smetal[i] * m;
        if (ag < 0.1) {
            a[3] += m;
            if (in50) a[4] += m;
            if (in90) a[5] += m;
        }
        if (ag < 0.01) {
            a[6] += m;
            if (in50) a[7] += m;
            if (in90) a[8] += m;
        }
        a[9] += m * svel[3 * i];
        a[10] += m * svel[3 * i + 1];
        a[11] += m * svel[3 * i + 2];
    }
    block_sums<12>(a, sh);
    const double msum = a[0];
    double *o = out + (int64_t)hlist[k] * S_NFIELDS;
    if (threadIdx.x == 0) {
        o[S_PX] = cx;
        o[S_PY] = cy;
        o[S_PZ] = cz;
        o[S_LX] = Lx;
        o[S_LY] = Ly;
        o[S_LZ] = Lz;
        o[S_RST] = sp[HP_R];
        o[S_R50] = r50;
        o[S_R90] = r90;
        o[S_AGE] = a[1] / msum;
        o[S_METAL] = a[2] / msum;
        o[S_SFR] = a[3] * 1.0e3;
        o[S_SFR50] = a[4] * 1.0e3;
        o[S_SFR90] = a[5] * 1.0e3;
        o[S_SFR10] = a[6] * 1.0e4;
        o[S_SFR10_50] = a[7] * 1.0e4;
        o[S_SFR10_90] = a[8] * 1.0e4;
    }
    if (!(n >= 2 && isfinite(msum) && msum > 0)) return;     // kinematics stay NaN
    const double bx = a[9] / msum, by = a[10] / msum, bz = a[11] / msum;
    const double amn = sqrt(Lx * Lx + Ly * Ly + Lz * Lz);
    const bool axis_ok = isfinite(amn) && amn > 0;
    const double ax = Lx / amn, ay = Ly / amn, az = Lz / amn;
    // per star: velocity about the bulk, cylindrical components when defined
    auto kin = [&](int64_t i, double &dvx, double &dvy, double &dvz, double &vr, double &vp,
                   double &vz, double &d2) {
        double dx, dy, dz;
        d2 = rel(i, dx, dy, dz);
        dvx = svel[3 * i] - bx;
        dvy = svel[3 * i + 1] - by;
        dvz = svel[3 * i + 2] - bz;
        if (!axis_ok) return false;
        const double zp = dx * ax + dy * ay + dz * az;
        const double px = dx - zp * ax, py = dy - zp * ay, pz = dz - zp * az;
        const double rp = sqrt(px * px + py * py + pz * pz);
        if (!(rp > DBL_EPS)) return false;
        const double inv = 1.0 / rp;
        const double ex = px * inv, ey = py * inv, ez = pz * inv;
        const double fx = ay * ez - az * ey, fy = az * ex - ax * ez, fz = ax * ey - ay * ex;
        vr = dvx * ex + dvy * ey + dvz * ez;
        vp = dvx * fx + dvy * fy + dvz * fz;
        vz = dvx * ax + dvy * ay + dvz * az;
        return true;
    };
    // pass 2, per selection (all, inside r50, inside r90): count, weight, weighted dv^2, and
    // over the cylindrically defined stars count, weight, weighted vphi, vrad, vzyl
    double b[24] = {0};
    for (int64_t i = f + threadIdx.x; i < e; i += BS) {
        double dvx, dvy, dvz, vr = 0, vp = 0, vz = 0, d2;
        const bool cyl = kin(i, dvx, dvy, dvz, vr, vp, vz, d2);
        const double m = smass[i];
        if (!(isfinite(m) && m > 0)) continue;
        const bool sel[3] = {true, d2 < r50s, d2 < r90s};
        const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
        for (int s = 0; s < 3; ++s) {
            if (!sel[s]) continue;
            double *q = b + 8 * s;
            q[0] += 1.0;
            q[1] += m;
            q[2] += dv2 * m;
            if (cyl) {
                q[3] += 1.0;
                q[4] += m;
                q[5] += vp * m;
                q[6] += vr * m;
                q[7] += vz * m;
            }
        }
    }
    block_sums<24>(b, sh);
    double mr[3], mp[3], mz[3];
    for (int s = 0; s < 3; ++s) {
        const double *q = b + 8 * s;
        mp[s] = q[5] / q[4];
        mr[s] = q[6] / q[4];
        mz[s] = q[7] / q[4];
    }
    // pass 3: cylindrical dispersion about the means
    double c3[3] = {0, 0, 0};
    for (int64_t i = f + threadIdx.x; i < e; i += BS) {
        double dvx, dvy, dvz, vr = 0, vp = 0, vz = 0, d2;
        const bool cyl = kin(i, dvx, dvy, dvz, vr, vp, vz, d2);
        const double m = smass[i];
        if (!cyl || !(isfinite(m) && m > 0)) continue;
        const bool sel[3] = {true, d2 < r50s, d2 < r90s};
        const double t[3] = {vr, vp, vz};
        for (int s = 0; s < 3; ++s) {
            if (!sel[s]) continue;
            const double a0 = t[0] - mr[s], a1 = t[1] - mp[s], a2 = t[2] - mz[s];
            c3[s] += (a0 * a0 + a1 * a1 + a2 * a2) * m;
        }
    }
    block_sums<3>(c3, sh);
    if (threadIdx.x == 0)
        for (int s = 0; s < 3; ++s) {
            const double *q = b + 8 * s;
            double vrot = NAN, sig3d = NAN, sigcyl = NAN;
            if (q[0] >= 2 && isfinite(q[1]) && q[1] > 0) {
                sig3d = sqrt(fmax(q[2] / q[1], 0.0));
                if (q[3] >= 2) {
                    vrot = mp[s];
                    sigcyl = sqrt(fmax(c3[s] / q[4], 0.0));
                }
            }
            o[S_VROT + 3 * s] = vrot;
            o[S_SIG3D + 3 * s] = sig3d;
            o[S_SIGCYL + 3 * s] = sigcyl;
        }
}

}  // namespace

float halo_stellar_device(const int64_t *h_off, int64_t nh, const double *pos, const double *vel,
                          const double *mass, const double *dens, const uint8_t *is_star,
                          const double *age, const double *metal, const double *chem,
                          int nchem, const double *props, const HaloParamsDev &P, double massp,
                          double *out, double *chem_out, hipStream_t s)
{
    if (nchem < 0 || nchem > MAX_CHEM || (nchem > 0 && (!chem || !chem_out)))
        throw std::runtime_error("nchem must be in [0, 9] with chemistry arrays");
    if (nh < 1 || nh > INT32_MAX / 2) throw std::runtime_error("need 1 <= nhalos < 2^30");
    if (h_off[0] != 0) throw std::runtime_error("offsets must start at 0");
    for (int64_t h = 0; h < nh; ++h)
        if (h_off[h + 1] <= h_off[h]) throw std::runtime_error("every halo needs at least one member");
    const int64_t M = h_off[nh];
    if (M > INT32_MAX) throw std::runtime_error("at most 2^31-1 members");
    hipEvent_t e0, e1;
    AHOP_HIP(hipEventCreate(&e0));
    AHOP_HIP(hipEventCreate(&e1));
    static DevBuf doff, dsc, dsoff;
    int64_t *d_off = (int64_t *)doff.ensure(sizeof(int64_t) * (nh + 1));
    int64_t *d_scount = (int64_t *)dsc.ensure(sizeof(int64_t) * nh);
    AHOP_HIP(hipMemcpyAsync(d_off, h_off, sizeof(int64_t) * (nh + 1), hipMemcpyHostToDevice, s));
    AHOP_HIP(hipEventRecord(e0, s));
    if (nchem > 0)   // halos without stars keep NaN (all-ones bytes)
        AHOP_HIP(hipMemsetAsync(chem_out, 0xFF, sizeof(double) * nchem * nh, s));
    // tiles of the type sums (the host lists live until the synchronisation below)
    std::vector<int32_t> th, tf(nh + 1);
    std::vector<int64_t> ts, te;
    {
        for (int64_t h = 0; h < nh; ++h) {
            tf[h] = (int32_t)th.size();
            for (int64_t a = h_off[h]; a < h_off[h + 1]; a += TYPE_TILE) {
                th.push_back((int32_t)h);
                ts.push_back(a);
                te.push_back(std::min(a + TYPE_TILE, h_off[h + 1]));
            }
        }
        tf[nh] = (int32_t)th.size();
        const int64_t nt = (int64_t)th.size();
        static DevBuf bth, btf, bts, bte, bpart;
        int32_t *d_th = (int32_t *)bth.ensure(sizeof(int32_t) * nt);
        int32_t *d_tf = (int32_t *)btf.ensure(sizeof(int32_t) * (nh + 1));
        int64_t *d_ts = (int64_t *)bts.ensure(sizeof(int64_t) * nt);
        int64_t *d_te = (int64_t *)bte.ensure(sizeof(int64_t) * nt);
        double *d_part = (double *)bpart.ensure(sizeof(double) * 7 * nt);
        AHOP_HIP(hipMemcpyAsync(d_th, th.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_tf, tf.data(), sizeof(int32_t) * (nh + 1), hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_ts, ts.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_te, te.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_type_sums, dim3((unsigned)nt), dim3(BS), 0, s, d_th, d_ts, d_te, vel,
                           mass, is_star, props, massp, d_part);
        hipLaunchKernelGGL(k_type_combine, dim3((unsigned)((nh + 127) / 128)), dim3(128), 0, s, nh,
                           d_tf, d_part, d_scount, out);
    }
    std::vector<int64_t> scount(nh);
    AHOP_HIP(hipMemcpyAsync(scount.data(), d_scount, sizeof(int64_t) * nh, hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> soff(nh + 1, 0), sub(1, 0);
    std::vector<int32_t> hlist;
    for (int64_t h = 0; h < nh; ++h) {
        soff[h + 1] = soff[h] + scount[h];
        if (scount[h] > 0) {
            hlist.push_back((int32_t)h);
            sub.push_back(soff[h + 1]);
        }
    }
    const int64_t S = soff[nh];
    const int nsh = (int)hlist.size();
    float ms = 0;
    if (S > 0) {
        if (!dens) throw std::runtime_error("the stellar centre needs densities");
        static DevBuf bsoff, bsub, bhl, bpos, bvel, bm, bd, ba, bz, bprops, br50, br90, bchem;
        double *schem = nchem > 0 ? (double *)bchem.ensure(sizeof(double) * nchem * S) : nullptr;
        int64_t *d_soff = (int64_t *)bsoff.ensure(sizeof(int64_t) * (nh + 1));
        int64_t *d_sub = (int64_t *)bsub.ensure(sizeof(int64_t) * (nsh + 1));
        int32_t *d_hl = (int32_t *)bhl.ensure(sizeof(int32_t) * nsh);
        double *spos = (double *)bpos.ensure(sizeof(double) * 3 * S);
        double *svel = (double *)bvel.ensure(sizeof(double) * 3 * S);
        double *smass = (double *)bm.ensure(sizeof(double) * S);
        double *sdens = (double *)bd.ensure(sizeof(double) * S);
        double *sage = (double *)ba.ensure(sizeof(double) * S);
        double *smetal = (double *)bz.ensure(sizeof(double) * S);
        double *sprops = (double *)bprops.ensure(sizeof(double) * PROPS_STRIDE * nsh);
        double *r50 = (double *)br50.ensure(sizeof(double) * nsh);
        double *r90 = (double *)br90.ensure(sizeof(double) * nsh);
        AHOP_HIP(hipMemcpyAsync(d_soff, soff.data(), sizeof(int64_t) * (nh + 1), hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_sub, sub.data(), sizeof(int64_t) * (nsh + 1), hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_hl, hlist.data(), sizeof(int32_t) * nsh, hipMemcpyHostToDevice, s));
        {
            static DevBuf bflag, bplace, bscan;
            int32_t *flag = (int32_t *)bflag.ensure(sizeof(int32_t) * M);
            int32_t *place = (int32_t *)bplace.ensure(sizeof(int32_t) * M);
            const unsigned gm = (unsigned)((M + 255) / 256);
            hipLaunchKernelGGL(k_star_flags, dim3(gm), dim3(256), 0, s, is_star, M, flag);
            size_t tb = 0;
            AHOP_HIP(rocprim::exclusive_scan(nullptr, tb, flag, place, 0, (size_t)M,
                                             rocprim::plus<int32_t>(), s));
            AHOP_HIP(rocprim::exclusive_scan(bscan.ensure(tb), tb, flag, place, 0, (size_t)M,
                                             rocprim::plus<int32_t>(), s));
            hipLaunchKernelGGL(k_star_scatter, dim3(gm), dim3(256), 0, s, is_star, place, M, pos,
                               vel, mass, dens, age, metal, chem, nchem, spos, svel, smass, sdens,
                               sage, smetal, schem);
        }
        // stellar centre (densest star, then DPMMC / SC over the stars), m*, L*, r*
        HaloParamsDev Q = P;
        Q.cdm = 0;
        Q.light = 1;
        (void)halo_props_device(sub.data(), nsh, spos, svel, smass, sdens, Q, sprops, 4096, 256, s);
        // stars of each halo in distance order: stable sorts by distance, then by star halo
        static DevBuf bdr, bdr2, bi, bi2, bh, bh2, bsd, bsm, temp;
        double *dr2 = (double *)bdr.ensure(sizeof(double) * S);
        double *dr2s = (double *)bdr2.ensure(sizeof(double) * S);
        int32_t *idx = (int32_t *)bi.ensure(sizeof(int32_t) * S);
        int32_t *idxs = (int32_t *)bi2.ensure(sizeof(int32_t) * S);
        uint32_t *hid = (uint32_t *)bh.ensure(sizeof(uint32_t) * S);
        uint32_t *hids = (uint32_t *)bh2.ensure(sizeof(uint32_t) * S);
        double *sd2 = (double *)bsd.ensure(sizeof(double) * S);
        double *sms = (double *)bsm.ensure(sizeof(double) * S);
        int hbits = 1;
        while (hbits < 32 && ((int64_t)(nsh - 1) >> hbits) != 0) ++hbits;
        hipLaunchKernelGGL(k_star_keys, dim3((unsigned)nsh), dim3(256), 0, s, d_sub, spos, sprops,
                           P.lbox, dr2, idx, hid);
        size_t tb1 = 0, tb2 = 0;
        AHOP_HIP(rocprim::radix_sort_pairs(nullptr, tb1, dr2, dr2s, idx, idxs, (size_t)S, 0, 64, s));
        AHOP_HIP(rocprim::radix_sort_pairs(nullptr, tb2, hids, hid, idxs, idx, (size_t)S, 0, hbits, s));
        void *tmp = temp.ensure(std::max(tb1, tb2));
        const unsigned gb = (unsigned)((S + 255) / 256);
        AHOP_HIP(rocprim::radix_sort_pairs(tmp, tb1, dr2, dr2s, idx, idxs, (size_t)S, 0, 64, s));
        hipLaunchKernelGGL(k_gather_u32, dim3(gb), dim3(256), 0, s, idxs, hid, (int64_t)S, hids);
        AHOP_HIP(rocprim::radix_sort_pairs(tmp, tb2, hids, hid, idxs, idx, (size_t)S, 0, hbits, s));
        hipLaunchKernelGGL(k_gather_sorted, dim3(gb), dim3(256), 0, s, idx, dr2, smass, (int64_t)S,
                           sd2, sms);
        static DevBuf bcum;
        double *cum = (double *)bcum.ensure(sizeof(double) * S);
        hipLaunchKernelGGL(k_star_r50, dim3((unsigned)nsh), dim3(BS), 0, s, d_sub, sd2, sms, cum,
                           r50, r90);
        hipLaunchKernelGGL(k_stellar, dim3((unsigned)nsh), dim3(BS), 0, s, d_hl, d_sub, spos, svel,
                           smass, sage, smetal, sprops, r50, r90, P.lbox, out);
        if (nchem > 0)
            hipLaunchKernelGGL(k_stellar_chem, dim3((unsigned)nsh), dim3(BS), 0, s, d_hl, d_sub,
                               smass, schem, nchem, chem_out);
        AHOP_HIP(hipStreamSynchronize(s));
    }
    // the span on the stream, host synchronisations for the star counts included
    AHOP_HIP(hipEventRecord(e1, s));
    AHOP_HIP(hipStreamSynchronize(s));
    AHOP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

}  // namespace ahop
