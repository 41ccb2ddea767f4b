// halo_props.hip -- batched per-halo property reductions on MI355X.
//
// Restates, for every halo of a CSR member layout at once, reference
// src/compute_halo_props.py:
//   det_mass_17 :754-780 (m), det_center_18 :901-966 (centre, mean velocity, sigma),
//   refine_centers_1b :1555-1567 (DPMMC :1447-1495 and SC :1498-1540 recentring),
//   compute_ang_mom_19 :783-824 (L), r_halos_1a :828-851 (r),
//   det_inertial_tensor_1b20 :1932-1951 + det_main_axis_1b2 :1954-1967 (I, axes),
//   det_halo_energies_1b3 :1400-1444 (ek, direct-sum ep below 1000 members),
//   tab_props_inside_1b40 :969-1032 + potential_profile_1b40 :1069-1123 (ellipsoidal
//   shell tables), det_vir_props_1b4 :1127-1229 (rvir, mvir, cvel, tvir),
//   compute_spin_parameter_1c :1599-1602 (spin).
//
// Work decomposition: halos are cut into tiles of at most T members (T picked by the
// autotuner in this file); every member pass is one launch with one workgroup per tile
// writing per-tile partials, combined per halo in tile order, so results do not depend
// on scheduling.  Accumulation is FP64.  Memory per member: pos 24 B, vel 24 B, mass 8 B
// read per pass.
#include "ahop_common.h"
#include "halo_props.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace ahop {

namespace {

constexpr int NSUM1 = 7;    // m, m*dr0 (3), m*v (3)
constexpr int NSUM3 = 12;   // m dv^2, L (3), I (6), m v2 (Hubble), r2max
constexpr double PI = 3.141592653589793;

struct HaloDev {
    const int64_t *off;
    const double *pos, *vel, *mass, *dens;
    const int32_t *tile_halo;
    const int64_t *tile_start, *tile_end;
    double L, hub, G;
};

__device__ __forceinline__ double wrap_strict(double d, double L)
{
    // correct_for_periodicity_1d :888-897
    if (d > 0.5 * L) d -= L;
    if (d < -0.5 * L) d += L;
    return d;
}

template <int BS, int NQ>
__device__ __forceinline__ void block_sum(double (&v)[NQ], double *sh)
{
    const int t = threadIdx.x;
    for (int q = 0; q < NQ; ++q) sh[q * BS + t] = v[q];
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < NQ; ++q) sh[q * BS + t] += sh[q * BS + t + s];
        __syncthreads();
    }
    for (int q = 0; q < NQ; ++q) v[q] = sh[q * BS];
    __syncthreads();
}

// ---- pass 1: mass, COM about the halo's first member, momentum
template <int BS>
__global__ __launch_bounds__(BS) void k_pass1(HaloDev D, double *part)
{
    __shared__ double sh[NSUM1 * BS];
    const int tile = blockIdx.x;
    const int h = D.tile_halo[tile];
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    const int64_t f = D.off[h];
    const double p0x = D.pos[3 * f], p0y = D.pos[3 * f + 1], p0z = D.pos[3 * f + 2];
    double v[NSUM1] = {0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        double m = D.mass[i];
        v[0] += m;
        v[1] += m * wrap_strict(D.pos[3 * i] - p0x, D.L);
        v[2] += m * wrap_strict(D.pos[3 * i + 1] - p0y, D.L);
        v[3] += m * wrap_strict(D.pos[3 * i + 2] - p0z, D.L);
        v[4] += m * D.vel[3 * i];
        v[5] += m * D.vel[3 * i + 1];
        v[6] += m * D.vel[3 * i + 2];
    }
    block_sum<BS, NSUM1>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < NSUM1; ++q) part[(int64_t)tile * NSUM1 + q] = v[q];
}

// per-halo state written by the combine kernels
struct HaloState {
    double m, pcx, pcy, pcz, vx, vy, vz;   // after pass 1
    double px, py, pz;                     // centre (after pass 2)
    double ax[3];                          // sha, shb, shc
    double ev[9];                          // eigenvectors, column j = axis j (row-major v[i][j])
    double amax, bmax, cmax;
    int64_t icenter;
};

__global__ void k_combine1(int nh, const int32_t *tile_first, const double *part,
                           const int64_t *off, const double *pos, double L, HaloState *st)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    double a[NSUM1] = {0, 0, 0, 0, 0, 0, 0};
    for (int t = tile_first[h]; t < tile_first[h + 1]; ++t)
        for (int q = 0; q < NSUM1; ++q) a[q] += part[(int64_t)t * NSUM1 + q];
    HaloState &S = st[h];
    int64_t f = off[h];
    double m = a[0];
    S.m = m;
    double c[3];
    for (int d = 0; d < 3; ++d) {
        double x = a[1 + d] / m + pos[3 * f + d];
        if (x > 0.5 * L) x -= L;             // correct_for_periodicity :869-876
        if (x < -0.5 * L) x += L;
        c[d] = x;
    }
    S.pcx = c[0];
    S.pcy = c[1];
    S.pcz = c[2];
    S.vx = a[4] / m;
    S.vy = a[5] / m;
    S.vz = a[6] / m;
}

// ---- pass 2: member nearest to the COM (cdm) or densest member; first index on ties
template <int BS>
__global__ __launch_bounds__(BS) void k_pass2(HaloDev D, const HaloState *st, int cdm,
                                              double *pkey, int64_t *pidx)
{
    __shared__ double sk[BS];
    __shared__ int64_t si[BS];
    const int tile = blockIdx.x;
    const int h = D.tile_halo[tile];
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    const HaloState &S = st[h];
    double best = INFINITY;
    int64_t bi = INT64_MAX;
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        double key;
        if (cdm) {
            double dx = wrap_strict(D.pos[3 * i] - S.pcx, D.L);
            double dy = wrap_strict(D.pos[3 * i + 1] - S.pcy, D.L);
            double dz = wrap_strict(D.pos[3 * i + 2] - S.pcz, D.L);
            key = dx * dx + dy * dy + dz * dz;
        } else {
            key = -D.dens[i];
        }
        if (key < best) {
            best = key;
            bi = i;
        }
    }
    const int t = threadIdx.x;
    sk[t] = best;
    si[t] = bi;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (t < w) {
            double k2 = sk[t + w];
            int64_t i2 = si[t + w];
            if (k2 < sk[t] || (k2 == sk[t] && i2 < si[t])) {
                sk[t] = k2;
                si[t] = i2;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        pkey[tile] = sk[0];
        pidx[tile] = si[0];
    }
}

__global__ void k_combine2(int nh, const int32_t *tile_first, const double *pkey,
                           const int64_t *pidx, const int64_t *off, const double *pos,
                           HaloState *st)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    double best = INFINITY;
    int64_t bi = -1;
    for (int t = tile_first[h]; t < tile_first[h + 1]; ++t)
        if (pkey[t] < best || (pkey[t] == best && pidx[t] < bi)) {
            best = pkey[t];
            bi = pidx[t];
        }
    if (bi < 0) bi = off[h];   // every key NaN (non-finite input): take the first member
    HaloState &S = st[h];
    S.icenter = bi;
    S.px = pos[3 * bi];
    S.py = pos[3 * bi + 1];
    S.pz = pos[3 * bi + 2];
}

// ---- recentring (refine_centers_1b :1555-1567): DPMMC, the multiscale grid
// (det_halo_center_multiscale_1b0 :1447-1495), then SC, the shrinking sphere
// (det_halo_center_sphere_1b1 :1498-1540).  One workgroup per halo; every grid level and
// every sphere is one block-wide pass over the halo's members (read from L2 after the first).
enum RecenterError { RC_DPMMC_INDEX = 1, RC_SC_DISTMIN = 2, RC_TOO_DEEP = 4 };
constexpr int64_t SC_TILED = 8192;     // SC alone: halos from this size on go sphere by sphere
                                       // over their member tiles
constexpr int RC_MAX_LEVELS = 256;     // grid levels (each shrinks the cube by 3)
constexpr int RC_MAX_SPHERES = 1000;   // spheres: the reference recurses once per sphere

// Block-wide reductions: lanes of a wave by cross-lane shuffles, then the waves in order
// through LDS (deterministic: a fixed combination order).  sh holds 2 * BS / 64 entries.
template <int BS>
__device__ double block_max_nan(double v, double *sh)
{
    // np.max: a NaN anywhere makes the maximum NaN
    auto mx = [](double a, double b) { return (a == a && (b != b || b > a)) ? b : a; };
    for (int o = 32; o > 0; o >>= 1) v = mx(v, __shfl_xor(v, o));
    if (BS > 64) {
        constexpr int NW = BS / 64;
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
        __syncthreads();
        v = sh[0];
        for (int w = 1; w < NW; ++w) v = mx(v, sh[w]);
        __syncthreads();
    }
    return v;
}

// smallest key, lowest index on ties (np.argmin's first occurrence)
template <int BS>
__device__ void block_argmin(double &key, int64_t &idx, double *sk, int64_t *si)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double k2 = __shfl_xor(key, o);
        const int64_t i2 = __shfl_xor(idx, o);
        if (k2 < key || (k2 == key && i2 < idx)) {
            key = k2;
            idx = i2;
        }
    }
    if (BS > 64) {
        constexpr int NW = BS / 64;
        if ((threadIdx.x & 63) == 0) {
            sk[threadIdx.x >> 6] = key;
            si[threadIdx.x >> 6] = idx;
        }
        __syncthreads();
        key = sk[0];
        idx = si[0];
        for (int w = 1; w < NW; ++w)
            if (sk[w] < key || (sk[w] == key && si[w] < idx)) {
                key = sk[w];
                idx = si[w];
            }
        __syncthreads();
    }
}

template <int BS, int NQ>
__device__ void block_sum_shfl(double (&v)[NQ], double *sh)
{
    for (int q = 0; q < NQ; ++q)
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
    if (BS > 64) {
        constexpr int NW = BS / 64;
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
}

// NGP cell of a position in an nx^3 grid from xmin with cells dx wide (:1460-1465): the
// quotient truncated toward zero must lie in [0, nx), i.e. the quotient in (-1, nx)
__device__ __forceinline__ bool ngp_cell(const double *p, const double *xmin, double dx, int nx,
                                         int &cell)
{
    int c[3];
    for (int d = 0; d < 3; ++d) {
        const double q = (p[d] - xmin[d]) / dx;
        if (!(q > -1.0 && q < (double)nx)) return false;
        c[d] = (int)q;
    }
    cell = (c[0] * nx + c[1]) * nx + c[2];
    return true;
}

// One workgroup per halo of the list; BS is sized to the halos' member counts by the host.
// grid: dynamic LDS of 9^3 doubles, present when dpmmc is on.
template <int BS>
__global__ __launch_bounds__(BS) void k_recenter(HaloDev D, HaloState *st, const int32_t *halos,
                                                 int sc, int dpmmc, double dcell_min,
                                                 double eps_sc, int *err)
{
    constexpr int NW = BS / 64;
    __shared__ double sh[4 * NW];
    __shared__ int64_t si[NW];
    __shared__ int scan[BS];
    __shared__ int64_t found;
    extern __shared__ double grid[];
    const int h = halos[blockIdx.x];
    const int t = threadIdx.x;
    const int64_t f = D.off[h], e = D.off[h + 1];
    const double L = D.L;
    HaloState &S = st[h];
    double c[3] = {S.px, S.py, S.pz};
    int64_t ic = S.icenter;

    // _radius_from_center :1543-1548
    auto radius = [&]() {
        double v = 0.0;
#pragma unroll 4
        for (int64_t i = f + t; i < e; i += BS) {
            const double dx = wrap_strict(D.pos[3 * i] - c[0], L);
            const double dy = wrap_strict(D.pos[3 * i + 1] - c[1], L);
            const double dz = wrap_strict(D.pos[3 * i + 2] - c[2], L);
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (v == v && (d2 != d2 || d2 > v)) v = d2;
        }
        return sqrt(block_max_nan<BS>(v, sh));
    };

    double r0 = radius();
    if (!(isfinite(r0) && r0 > 0.0)) return;       // block-uniform

    if (dpmmc) {
        // the densest member (first on ties): its index within the halo picks the centre
        double bk = INFINITY;
        int64_t bi = INT64_MAX;
        for (int64_t i = f + t; i < e; i += BS)
            if (-D.dens[i] < bk) {
                bk = -D.dens[i];
                bi = i;
            }
        block_argmin<BS>(bk, bi, sh, si);
        const int64_t a = bi == INT64_MAX ? 0 : bi - f;
        double x0[3] = {c[0], c[1], c[2]};
        double rr = r0;
        int nx = 9;
        int64_t target = -2;                           // -2: still descending
        for (int level = 0; level < RC_MAX_LEVELS; ++level) {
            const double dx = 2 * rr / nx;
            const double xmin[3] = {x0[0] - rr, x0[1] - rr, x0[2] - rr};
            const int ncell = nx * nx * nx;
            for (int q = t; q < ncell; q += BS) grid[q] = 0.0;
            __syncthreads();
            for (int64_t i = f + t; i < e; i += BS) {
                int cell;
                if (ngp_cell(D.pos + 3 * i, xmin, dx, nx, cell)) atomicAdd(&grid[cell], D.mass[i]);
            }
            __syncthreads();
            // first cell of maximum mass in (i, j, k) order (np.where's first, :1475-1477)
            double ck = INFINITY;
            int64_t cb = INT64_MAX;
            for (int q = t; q < ncell; q += BS)
                if (-grid[q] < ck) {
                    ck = -grid[q];
                    cb = q;
                }
            block_argmin<BS>(ck, cb, sh, si);
            if (dx > 3.0 * dcell_min) {
                const int cell = cb == INT64_MAX ? 0 : (int)cb;
                const int cx = cell / (nx * nx), cy = (cell / nx) % nx, cz = cell % nx;
                x0[0] = xmin[0] + (cx + 0.5) * dx;
                x0[1] = xmin[1] + (cy + 0.5) * dx;
                x0[2] = xmin[2] + (cz + 0.5) * dx;
                rr = dx;
                nx = 3;
                continue;
            }
            // member number a among the members inside this grid, in member order (:1486-1490)
            target = -1;
            int64_t seen = 0;
            for (int64_t base = f; base < e; base += BS) {
                const int64_t i = base + t;
                int cell;
                const int in = (i < e) && ngp_cell(D.pos + 3 * i, xmin, dx, nx, cell);
                if (t == 0) found = -1;
                scan[t] = in;
                __syncthreads();
                for (int w = 1; w < BS; w <<= 1) {
                    const int add = t >= w ? scan[t - w] : 0;
                    __syncthreads();
                    scan[t] += add;
                    __syncthreads();
                }
                if (in && seen + scan[t] - 1 == a) found = i;
                __syncthreads();
                seen += scan[BS - 1];
                const int64_t fi = found;
                __syncthreads();
                if (fi >= 0) {
                    target = fi;
                    break;
                }
            }
            break;
        }
        if (target < 0) {
            if (t == 0) atomicOr(err, target == -1 ? RC_DPMMC_INDEX : RC_TOO_DEEP);
            return;
        }
        ic = target;
        c[0] = D.pos[3 * target];
        c[1] = D.pos[3 * target + 1];
        c[2] = D.pos[3 * target + 2];
        r0 = radius();
    }

    if (sc) {
        double x0[3] = {c[0], c[1], c[2]};
        double r = r0;
        int64_t target = -1;
        for (int it = 0; it < RC_MAX_SPHERES; ++it) {
            const double r2 = r * r;
            double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
            for (int64_t i = f + t; i < e; i += BS) {
                const double dx = wrap_strict(D.pos[3 * i] - x0[0], L);
                const double dy = wrap_strict(D.pos[3 * i + 1] - x0[1], L);
                const double dz = wrap_strict(D.pos[3 * i + 2] - x0[2], L);
                if (dx * dx + dy * dy + dz * dz <= r2) {
                    const double m = D.mass[i];
                    v[0] += m;
                    v[1] += m * dx;
                    v[2] += m * dy;
                    v[3] += m * dz;
                }
            }
            block_sum_shfl<BS, 4>(v, sh);
            double xc[3] = {x0[0], x0[1], x0[2]};
            if (v[0] > 0.0)
                for (int d = 0; d < 3; ++d) xc[d] = v[1 + d] / v[0] + x0[d];
            if (r > dcell_min && v[0] > 0.0) {
                for (int d = 0; d < 3; ++d) x0[d] = xc[d];
                r = (1 - eps_sc) * r;
                continue;
            }
            // the member nearest the last centre of mass (first on ties)
            double bk = INFINITY;
            int64_t bi = INT64_MAX;
#pragma unroll 4
            for (int64_t i = f + t; i < e; i += BS) {
                const double dx = wrap_strict(D.pos[3 * i] - xc[0], L);
                const double dy = wrap_strict(D.pos[3 * i + 1] - xc[1], L);
                const double dz = wrap_strict(D.pos[3 * i + 2] - xc[2], L);
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < bk) {
                    bk = d2;
                    bi = i;
                }
            }
            block_argmin<BS>(bk, bi, sh, si);
            if (bi == INT64_MAX) bi = f;                // every distance NaN: np.argmin gives 0
            if (bk > 1.0) {
                if (t == 0) atomicOr(err, RC_SC_DISTMIN);
                return;
            }
            target = bi;
            break;
        }
        if (target < 0) {
            if (t == 0) atomicOr(err, RC_TOO_DEEP);
            return;
        }
        ic = target;
        c[0] = D.pos[3 * target];
        c[1] = D.pos[3 * target + 1];
        c[2] = D.pos[3 * target + 2];
    }
    if (t == 0) {
        S.px = c[0];
        S.py = c[1];
        S.pz = c[2];
        S.icenter = ic;
    }
}

// ---- SC for the largest halos, one sphere at a time over the halo's member tiles: every
// sphere is a partial-sum launch over the tiles and a per-halo combine in tile order.  The
// host replays the radius sequence (r0, (1-eps) r0, ...) to know how many spheres to launch.
struct ScState {
    double x[3];   // current centre; after the last sphere, its centre of mass
    double r;      // current radius
    int phase;     // 0: shrinking, 1: last centre of mass in x, 2: no recentring
};

template <int BS>
__global__ __launch_bounds__(BS) void k_sc_radius_tiles(HaloDev D, const HaloState *st,
                                                        const int32_t *tiles, double *pmax)
{
    __shared__ double sh[BS / 64];
    const int tile = tiles[blockIdx.x];
    const int h = D.tile_halo[tile];
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    const HaloState &S = st[h];
    double v = 0.0;
#pragma unroll 4
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        const double dx = wrap_strict(D.pos[3 * i] - S.px, D.L);
        const double dy = wrap_strict(D.pos[3 * i + 1] - S.py, D.L);
        const double dz = wrap_strict(D.pos[3 * i + 2] - S.pz, D.L);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (v == v && (d2 != d2 || d2 > v)) v = d2;
    }
    v = block_max_nan<BS>(v, sh);
    if (threadIdx.x == 0) pmax[tile] = v;
}

__global__ void k_sc_init(const int32_t *halos, int nb, const int32_t *tile_first,
                          const double *pmax, const HaloState *st, ScState *sc, double *r0_out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const int h = halos[k];
    double m = pmax[tile_first[h]];
    for (int t = tile_first[h] + 1; t < tile_first[h + 1]; ++t) {
        const double b = pmax[t];
        if (m == m && (b != b || b > m)) m = b;
    }
    const double r0 = sqrt(m);
    ScState &C = sc[h];
    C.x[0] = st[h].px;
    C.x[1] = st[h].py;
    C.x[2] = st[h].pz;
    C.r = r0;
    C.phase = (isfinite(r0) && r0 > 0.0) ? 0 : 2;
    r0_out[k] = r0;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_sc_sphere_tiles(HaloDev D, const ScState *sc,
                                                        const int32_t *tiles, double *part)
{
    __shared__ double sh[4 * (BS / 64)];
    const int tile = tiles[blockIdx.x];
    const int h = D.tile_halo[tile];
    const ScState &C = sc[h];
    if (C.phase != 0) return;                      // block-uniform
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    const double r2 = C.r * C.r;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        const double dx = wrap_strict(D.pos[3 * i] - C.x[0], D.L);
        const double dy = wrap_strict(D.pos[3 * i + 1] - C.x[1], D.L);
        const double dz = wrap_strict(D.pos[3 * i + 2] - C.x[2], D.L);
        if (dx * dx + dy * dy + dz * dz <= r2) {
            const double m = D.mass[i];
            v[0] += m;
            v[1] += m * dx;
            v[2] += m * dy;
            v[3] += m * dz;
        }
    }
    block_sum_shfl<BS, 4>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) part[(int64_t)tile * 4 + q] = v[q];
}

__global__ void k_sc_sphere_combine(const int32_t *halos, int nb, const int32_t *tile_first,
                                    const double *part, ScState *sc, double dcell_min,
                                    double eps_sc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const int h = halos[k];
    ScState &C = sc[h];
    if (C.phase != 0) return;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = tile_first[h]; t < tile_first[h + 1]; ++t)
        for (int q = 0; q < 4; ++q) a[q] += part[(int64_t)t * 4 + q];
    double xc[3] = {C.x[0], C.x[1], C.x[2]};
    if (a[0] > 0.0)
        for (int d = 0; d < 3; ++d) xc[d] = a[1 + d] / a[0] + C.x[d];
    for (int d = 0; d < 3; ++d) C.x[d] = xc[d];
    if (C.r > dcell_min && a[0] > 0.0)
        C.r = (1 - eps_sc) * C.r;
    else
        C.phase = 1;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_sc_nearest_tiles(HaloDev D, const ScState *sc,
                                                         const int32_t *tiles, double *pkey,
                                                         int64_t *pidx)
{
    __shared__ double sk[BS / 64];
    __shared__ int64_t si[BS / 64];
    const int tile = tiles[blockIdx.x];
    const int h = D.tile_halo[tile];
    const ScState &C = sc[h];
    if (C.phase != 1) return;                      // block-uniform
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    double bk = INFINITY;
    int64_t bi = INT64_MAX;
#pragma unroll 4
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        const double dx = wrap_strict(D.pos[3 * i] - C.x[0], D.L);
        const double dy = wrap_strict(D.pos[3 * i + 1] - C.x[1], D.L);
        const double dz = wrap_strict(D.pos[3 * i + 2] - C.x[2], D.L);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < bk) {
            bk = d2;
            bi = i;
        }
    }
    block_argmin<BS>(bk, bi, sk, si);
    if (threadIdx.x == 0) {
        pkey[tile] = bk;
        pidx[tile] = bi;
    }
}

__global__ void k_sc_nearest_combine(const int32_t *halos, int nb, const int32_t *tile_first,
                                     const double *pkey, const int64_t *pidx, const int64_t *off,
                                     const double *pos, const ScState *sc, HaloState *st,
                                     int *err)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const int h = halos[k];
    const ScState &C = sc[h];
    if (C.phase == 2) return;
    if (C.phase == 0) {                            // more spheres than the host launched
        atomicOr(err, RC_TOO_DEEP);
        return;
    }
    double best = INFINITY;
    int64_t bi = INT64_MAX;
    for (int t = tile_first[h]; t < tile_first[h + 1]; ++t)
        if (pkey[t] < best || (pkey[t] == best && pidx[t] < bi)) {
            best = pkey[t];
            bi = pidx[t];
        }
    if (bi == INT64_MAX) bi = off[h];              // every distance NaN: np.argmin gives 0
    if (best > 1.0) {
        atomicOr(err, RC_SC_DISTMIN);
        return;
    }
    HaloState &S = st[h];
    S.icenter = bi;
    S.px = pos[3 * bi];
    S.py = pos[3 * bi + 1];
    S.pz = pos[3 * bi + 2];
}

// ---- pass 3: second moments about the centre
template <int BS>
__global__ __launch_bounds__(BS) void k_pass3(HaloDev D, const HaloState *st, double *part)
{
    __shared__ double sh[NSUM3 * BS];
    const int tile = blockIdx.x;
    const int h = D.tile_halo[tile];
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    const HaloState &S = st[h];
    double v[NSUM3] = {0};
    double r2max = 0.0;
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        double m = D.mass[i];
        double dx = wrap_strict(D.pos[3 * i] - S.px, D.L);
        double dy = wrap_strict(D.pos[3 * i + 1] - S.py, D.L);
        double dz = wrap_strict(D.pos[3 * i + 2] - S.pz, D.L);
        double ux = D.vel[3 * i] - S.vx, uy = D.vel[3 * i + 1] - S.vy, uz = D.vel[3 * i + 2] - S.vz;
        v[0] += m * (ux * ux + uy * uy + uz * uz);
        double pvx = m * ux, pvy = m * uy, pvz = m * uz;
        v[1] += dy * pvz - dz * pvy;
        v[2] += dz * pvx - dx * pvz;
        v[3] += dx * pvy - dy * pvx;
        v[4] += m * dx * dx;
        v[5] += m * dx * dy;
        v[6] += m * dx * dz;
        v[7] += m * dy * dy;
        v[8] += m * dy * dz;
        v[9] += m * dz * dz;
        double tx = ux + dx * D.hub, ty = uy + dy * D.hub, tz = uz + dz * D.hub;
        v[10] += m * (tx * tx + ty * ty + tz * tz);
        r2max = fmax(r2max, dx * dx + dy * dy + dz * dz);
    }
    v[11] = r2max;
    // sums for 0..10, max for 11
    const int t = threadIdx.x;
    for (int q = 0; q < NSUM3; ++q) sh[q * BS + t] = v[q];
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int q = 0; q < NSUM3 - 1; ++q) sh[q * BS + t] += sh[q * BS + t + w];
            sh[11 * BS + t] = fmax(sh[11 * BS + t], sh[11 * BS + t + w]);
        }
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < NSUM3; ++q) part[(int64_t)tile * NSUM3 + q] = sh[q * BS];
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (FP64), eigenvalues
// ascending with matching eigenvector columns, as scipy.linalg.eigh returns them.
__device__ void eigh3(const double A[6], double d[3], double V[9])
{
    double a[3][3] = {{A[0], A[1], A[2]}, {A[1], A[3], A[4]}, {A[2], A[4], A[5]}};
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (off <= 1e-34 * diag || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = a[p][q];
                if (apq == 0.0) continue;
                double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {   // A <- A J
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {   // A <- J^T A
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int idx[3] = {0, 1, 2};
    double w[3] = {a[0][0], a[1][1], a[2][2]};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2 - i; ++j)
            if (w[idx[j]] > w[idx[j + 1]]) {
                int tmp = idx[j];
                idx[j] = idx[j + 1];
                idx[j + 1] = tmp;
            }
    for (int j = 0; j < 3; ++j) {
        d[j] = w[idx[j]];
        for (int i = 0; i < 3; ++i) V[3 * i + j] = v[i][idx[j]];
    }
}

__global__ void k_combine3(int nh, const int32_t *tile_first, const double *part, double G,
                           HaloState *st, double *out)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    double a[NSUM3] = {0};
    for (int t = tile_first[h]; t < tile_first[h + 1]; ++t) {
        for (int q = 0; q < NSUM3 - 1; ++q) a[q] += part[(int64_t)t * NSUM3 + q];
        a[11] = fmax(a[11], part[(int64_t)t * NSUM3 + 11]);
    }
    HaloState &S = st[h];
    double *o = out + (int64_t)h * HP_NFIELDS;
    o[HP_M] = S.m;
    o[HP_PX] = S.px;
    o[HP_PY] = S.py;
    o[HP_PZ] = S.pz;
    o[HP_VX] = S.vx;
    o[HP_VY] = S.vy;
    o[HP_VZ] = S.vz;
    o[HP_SIGMA] = sqrt(a[0] / S.m);
    o[HP_LX] = a[1];
    o[HP_LY] = a[2];
    o[HP_LZ] = a[3];
    o[HP_R] = sqrt(a[11]);
    for (int q = 0; q < 6; ++q) o[HP_IXX + q] = a[4 + q];
    o[HP_EK] = 0.5 * a[10];
    double d[3];
    eigh3(&a[4], d, S.ev);
    for (int j = 0; j < 3; ++j) {
        S.ax[j] = sqrt(d[j] / S.m);
        o[HP_SHA + j] = S.ax[j];
    }
}

__device__ __forceinline__ void ell_proj(const HaloDev &D, const HaloState &S, int64_t i,
                                         double &ra, double &rb, double &rc)
{
    double dx = wrap_strict(D.pos[3 * i] - S.px, D.L);
    double dy = wrap_strict(D.pos[3 * i + 1] - S.py, D.L);
    double dz = wrap_strict(D.pos[3 * i + 2] - S.pz, D.L);
    // dras = drx*v[0,0] + dry*v[1,0] + drz*v[2,0] (tab_props_inside_1b40 :984-986)
    ra = dx * S.ev[0] + dy * S.ev[3] + dz * S.ev[6];
    rb = dx * S.ev[1] + dy * S.ev[4] + dz * S.ev[7];
    rc = dx * S.ev[2] + dy * S.ev[5] + dz * S.ev[8];
}

// ---- pass 4a: largest ellipsoidal radius in units of the principal axes
template <int BS>
__global__ __launch_bounds__(BS) void k_pass4a(HaloDev D, const HaloState *st, double *pmax)
{
    __shared__ double sh[BS];
    const int tile = blockIdx.x;
    const int h = D.tile_halo[tile];
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    const HaloState &S = st[h];
    double mx = 0.0;
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        double ra, rb, rc;
        ell_proj(D, S, i, ra, rb, rc);
        double qa = ra / S.ax[0], qb = rb / S.ax[1], qc = rc / S.ax[2];
        mx = fmax(mx, sqrt(qa * qa + qb * qb + qc * qc));
    }
    const int t = threadIdx.x;
    sh[t] = mx;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] = fmax(sh[t], sh[t + w]);
        __syncthreads();
    }
    if (t == 0) pmax[tile] = sh[0];
}

__global__ void k_combine4a(int nh, const int32_t *tile_first, const double *pmax, HaloState *st)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    double rmax = 0.0;
    for (int t = tile_first[h]; t < tile_first[h + 1]; ++t) rmax = fmax(rmax, pmax[t]);
    HaloState &S = st[h];
    const double eps = 1e-2;
    S.amax = rmax * S.ax[0] * (1.0 + eps);
    S.bmax = rmax * S.ax[1] * (1.0 + eps);
    S.cmax = rmax * S.ax[2] * (1.0 + eps);
}

// ---- pass 4b: per-tile shell histograms (mass, kinetic energy, count) in LDS
template <int BS>
__global__ __launch_bounds__(BS) void k_pass4b(HaloDev D, const HaloState *st, int nr, double *hist,
                                               int *missed)
{
    extern __shared__ double hs[];   // 3*nr doubles
    const int tile = blockIdx.x;
    const int h = D.tile_halo[tile];
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    const HaloState &S = st[h];
    for (int b = threadIdx.x; b < 3 * nr; b += BS) hs[b] = 0.0;
    __syncthreads();
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        double ra, rb, rc;
        ell_proj(D, S, i, ra, rb, rc);
        double qa = ra / S.amax, qb = rb / S.bmax, qc = rc / S.cmax;
        double r = sqrt(qa * qa + qb * qb + qc * qc);
        double fi = r * nr;
        int ib = (int)fi;
        if (!(fi < (double)nr) || ib >= nr) {
            atomicAdd(missed, 1);
            continue;
        }
        double m = D.mass[i];
        double dx = wrap_strict(D.pos[3 * i] - S.px, D.L);
        double dy = wrap_strict(D.pos[3 * i + 1] - S.py, D.L);
        double dz = wrap_strict(D.pos[3 * i + 2] - S.pz, D.L);
        double tx = D.vel[3 * i] - S.vx + dx * D.hub;
        double ty = D.vel[3 * i + 1] - S.vy + dy * D.hub;
        double tz = D.vel[3 * i + 2] - S.vz + dz * D.hub;
        double v2 = tx * tx + ty * ty + tz * tz;
        atomicAdd(&hs[ib], m);
        atomicAdd(&hs[nr + ib], 0.5 * m * v2);
        atomicAdd(&hs[2 * nr + ib], 1.0);
    }
    __syncthreads();
    double *out = hist + (int64_t)tile * 3 * nr;
    for (int b = threadIdx.x; b < 3 * nr; b += BS) out[b] = hs[b];
}

// ---- per-shell mass and angular momentum (det_ang_momentum_per_shell_1b5 :1970-2013), the
// ANG_MOM_OF_R output: members binned into nsh ellipsoidal shells of the all-member ellipsoid
// (amax, bmax, cmax of tab_props_inside_1b40), i_ell = int(r_ell * nsh); per shell the mass
// and L = sum m (dr x dv), dr the wrapped offset from the centre, dv = v - the halo's mean
// velocity (no Hubble term).  Per-tile partials in LDS (4*nsh doubles), combined per halo in
// tile order by k_angmom_combine.
template <int BS>
__global__ __launch_bounds__(BS) void k_angmom(HaloDev D, const HaloState *st, int nsh,
                                               double *hist, int *missed)
{
    extern __shared__ double hs[];   // 4*nsh doubles: m, Lx, Ly, Lz
    const int tile = blockIdx.x;
    const int h = D.tile_halo[tile];
    const int64_t s = D.tile_start[tile], e = D.tile_end[tile];
    const HaloState &S = st[h];
    for (int b = threadIdx.x; b < 4 * nsh; b += BS) hs[b] = 0.0;
    __syncthreads();
    for (int64_t i = s + threadIdx.x; i < e; i += BS) {
        double ra, rb, rc;
        ell_proj(D, S, i, ra, rb, rc);
        // r_ell = sqrt((dra / amax)**2 + (drb / bmax)**2 + (drc / cmax)**2) (:1988)
        const double qa = ra / S.amax, qb = rb / S.bmax, qc = rc / S.cmax;
        const double fi = sqrt(qa * qa + qb * qb + qc * qc) * nsh;
        const int ib = (int)fi;
        if (!(fi < (double)nsh) || ib >= nsh) {
            atomicAdd(missed, 1);
            continue;
        }
        const double m = D.mass[i];
        const double dx = wrap_strict(D.pos[3 * i] - S.px, D.L);
        const double dy = wrap_strict(D.pos[3 * i + 1] - S.py, D.L);
        const double dz = wrap_strict(D.pos[3 * i + 2] - S.pz, D.L);
        const double vx = D.vel[3 * i] - S.vx, vy = D.vel[3 * i + 1] - S.vy,
                     vz = D.vel[3 * i + 2] - S.vz;
        atomicAdd(&hs[ib], m);
        atomicAdd(&hs[nsh + ib], m * (dy * vz - dz * vy));
        atomicAdd(&hs[2 * nsh + ib], m * (dz * vx - dx * vz));
        atomicAdd(&hs[3 * nsh + ib], m * (dx * vy - dy * vx));
    }
    __syncthreads();
    double *o = hist + (int64_t)tile * 4 * nsh;
    for (int b = threadIdx.x; b < 4 * nsh; b += BS) o[b] = hs[b];
}

// per halo: its tiles' shell sums in tile order into shells (nh x 4 x nsh: m, Lx, Ly, Lz per
// shell) and the ellipsoid (amax, bmax, cmax, then the eigenvectors v row-major) into axes
// (nh x 12).  The reference writes a halo's record only when its rvir is finite (:1257);
// the other halos' rows are NaN.
__global__ void k_angmom_combine(int nh, const int32_t *tile_first, const double *hist, int nsh,
                                 const HaloState *st, const double *props, double *shells,
                                 double *axes)
{
    const int h = blockIdx.x;
    if (h >= nh) return;
    const bool keep = isfinite(props[(int64_t)h * HP_NFIELDS + HP_RVIR]);
    const int t0 = tile_first[h], t1 = tile_first[h + 1];
    for (int b = threadIdx.x; b < 4 * nsh; b += blockDim.x) {
        double acc = 0.0;
        for (int t = t0; t < t1; ++t) acc += hist[(int64_t)t * 4 * nsh + b];
        shells[(int64_t)h * 4 * nsh + b] = keep ? acc : NAN;
    }
    if (threadIdx.x < 12) {
        const HaloState &S = st[h];
        const int q = threadIdx.x;
        const double v = q == 0 ? S.amax : q == 1 ? S.bmax : q == 2 ? S.cmax : S.ev[q - 3];
        axes[(int64_t)h * 12 + q] = keep ? v : NAN;
    }
}

// 1/sqrt(d) from v_rsq_f64 (relative error below 2^-24) and one Newton step: below 5e-15
// relative over 2^-40..2^5 (measured on gfx950 against 1/sqrt in long double), far inside the
// pair sum's 1e-10.  d = 0 (members at one position) gives NaN, not +inf: such a halo's sum is
// redone with the library's rsqrt.
__device__ __forceinline__ double rsqrt_nr1(double d)
{
    const double y0 = __builtin_amdgcn_rsq(d);
    const double e = fma(-(d * y0), y0, 1.0);
    return fma(0.5 * y0, e, y0);
}

// sum over members j0 <= j < je of m_j / d_ij for member a = mem[i] (for a tile on the
// diagonal only the j > i): j is the same on every lane of the wave (the member is read
// once for all 64 lanes), four partial sums carry independent rsqrt chains.  FOLD takes the
// minimum image min(|t|, L - |t|) per axis.  The distance is formed with fused multiply-adds
// and 1/d with rsqrt_nr1 (EXACT: the library's rsqrt): the sum is held to 1e-10 of the
// energies, not bit for bit.
template <bool FOLD, bool DIAG, bool EXACT>
__device__ __forceinline__ double pe_tile(const double4 *mem, const double4 a, int i, int j0,
                                          int je, double L)
{
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    auto term = [&](int j, double &racc) {
        const double4 b = mem[j];
        double tx = b.x - a.x, ty = b.y - a.y, tz = b.z - a.z;
        if (FOLD) {
            tx = fabs(tx);
            ty = fabs(ty);
            tz = fabs(tz);
            tx = fmin(tx, L - tx);
            ty = fmin(ty, L - ty);
            tz = fmin(tz, L - tz);
        }
        double d2 = fma(tz, tz, fma(ty, ty, tx * tx));
        double w = b.w;
        if (DIAG) {
            const bool on = j > i;
            d2 = on ? d2 : 1.0;
            w = on ? w : 0.0;
        }
        racc = fma(w, EXACT ? rsqrt(d2) : rsqrt_nr1(d2), racc);
    };
    int j = j0;
    for (; j + 4 <= je; j += 4) {
        term(j, r0);
        term(j + 1, r1);
        term(j + 2, r2);
        term(j + 3, r3);
    }
    for (; j < je; ++j) term(j, r0);
    return (r0 + r1) + (r2 + r3);
}

// ---- direct potential energy of halos below the pair-sum limit (det_halo_energies_1b3
// :1417-1427): ep = sum_{i<j} -(m_i m_j) / (d_ij / L) * G / L = -G sum m_i m_j / d_ij, with
// d_ij the minimum-image distance (the strict wrap at L/2 gives the same distance as
// min(|d|, L - |d|)).  The halo's members are staged in LDS (at most PE_NMAX of them) and cut
// into tiles of 64; the T(T+1)/2 tile pairs (A <= B) go round the block's waves.  In a tile
// pair lane l holds member 64A + l and runs over the members of tile B in step with the
// other lanes (on the diagonal A = B only the pairs j > i count).
constexpr int PE_NMAX = 2048;   // 64 KiB of LDS
constexpr int PE_TILE = 64;     // members per tile: one per lane of a wave
template <int BS, bool EXACT>
__global__ __launch_bounds__(BS) void k_direct_pe(HaloDev D, const int32_t *small_halos,
                                                  double *ep_out)
{
    static_assert(BS % PE_TILE == 0, "whole waves");
    extern __shared__ double4 mem[];
    __shared__ double sh[BS];
    const int h = small_halos[blockIdx.x];
    // the exact pass redoes only the halos whose fast sum is NaN
    if (EXACT && !isnan(ep_out[h])) return;
    const int64_t f = D.off[h];
    const int n = (int)(D.off[h + 1] - f);
    for (int i = threadIdx.x; i < n; i += BS)
        mem[i] = make_double4(D.pos[3 * (f + i)], D.pos[3 * (f + i) + 1], D.pos[3 * (f + i) + 2],
                              D.mass[f + i]);
    const double L = D.L;
    // a halo whose members all lie within L/4 of its first member on every axis has every
    // pair closer than L/2 per axis, where min(t, L - t) = t: its pairs skip the fold
    bool far = false;
    {
        const double4 m0 = mem[0];
        for (int i = threadIdx.x; i < n; i += BS) {
            const double4 c = mem[i];
            far = far || fabs(c.x - m0.x) >= 0.25 * L || fabs(c.y - m0.y) >= 0.25 * L ||
                  fabs(c.z - m0.z) >= 0.25 * L;
        }
    }
    const bool fold = __syncthreads_or(far);
    const int lane = threadIdx.x % PE_TILE;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / PE_TILE);
    const int T = (n + PE_TILE - 1) / PE_TILE;
    const int NT = T * (T + 1) / 2;
    double acc = 0.0;
    for (int q = wave; q < NT; q += BS / PE_TILE) {
        int A = 0, rem = q;   // tile pair q, numbered row by row: (A, A + rem)
        while (rem >= T - A) {
            rem -= T - A;
            ++A;
        }
        const int B = A + rem;
        const int i = A * PE_TILE + lane;
        const int j0 = B * PE_TILE, je = min(j0 + PE_TILE, n);
        if (i < n) {
            const double4 a = mem[i];
            double row;
            if (fold)
                row = A == B ? pe_tile<true, true, EXACT>(mem, a, i, j0, je, L)
                             : pe_tile<true, false, EXACT>(mem, a, i, j0, je, L);
            else
                row = A == B ? pe_tile<false, true, EXACT>(mem, a, i, j0, je, L)
                             : pe_tile<false, false, EXACT>(mem, a, i, j0, je, L);
            acc = fma(a.w, row, acc);
        }
    }
    const int t = threadIdx.x;
    sh[t] = acc;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) ep_out[h] = -D.G * sh[0];
}

// In-place inclusive prefix sum added strictly left to right (np.cumsum's order), one thread;
// loads are batched so that only the additions form the dependent chain.
__device__ void seq_cumsum(double *a, int n)
{
    double acc = 0.0;
    int b = 0;
    for (; b + 8 <= n; b += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = a[b + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            acc += v[u];
            a[b + u] = acc;
        }
    }
    for (; b < n; ++b) {
        acc += a[b];
        a[b] = acc;
    }
}

constexpr int WAVE_HP = 64;   // first thread of the second wave of a workgroup

// Carlson's symmetric elliptic integral RF(x, y, z) by duplication (B. C. Carlson, "Numerical
// computation of real or complex elliptic integrals", Numer. Algorithms 10 (1995) 13-26,
// eqs. 2.2-2.4): the arguments are pulled together until the fifth-order series about their
// mean holds to double precision.  The reference evaluates it with scipy.special.elliprf.
__device__ double carlson_rf(double x, double y, double z)
{
    double a = (x + y + z) / 3.0;
    const double a0 = a, x0 = x, y0 = y;
    // (3 r)^(-1/6) with r = 2^-53 bounds the truncation of the series at r
    double q = 4.4e2 * fmax(fabs(a - x), fmax(fabs(a - y), fabs(a - z)));
    double f = 1.0;   // 4^-n
    for (int n = 0; n < 64 && q * f >= fabs(a); ++n) {
        const double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        const double lam = sx * sy + sx * sz + sy * sz;
        x = 0.25 * (x + lam);
        y = 0.25 * (y + lam);
        z = 0.25 * (z + lam);
        a = 0.25 * (a + lam);
        f *= 0.25;
    }
    const double X = (a0 - x0) * f / a, Y = (a0 - y0) * f / a, Z = -(X + Y);
    const double e2 = X * Y - Z * Z, e3 = X * Y * Z;
    return (1.0 - e2 / 10.0 + e3 / 14.0 + e2 * e2 / 24.0 - 3.0 * e2 * e3 / 44.0) / sqrt(a);
}

// ---- the virial radius and mass from the scans over the shell tables (det_vir_props_1b4
// :1127-1229) and the output fields.  The scans pick the outermost shell meeting a condition:
//   i1r -- outermost shell holding mass (tm[i-1] < tm[i]), 0 if none;
//   brk -- outermost bound shell with ratio <= 0.2 (-1 if none): the virial scan stops there;
//          that shell is always a new best, so it sets mvir (tm_brk) and (the axes being a
//          running minimum of increasing functions of the index) the axes; without such a
//          shell the last ratio seen exceeds 0.2 and the overdensity branch decides;
//   fb  -- outermost shell with tm >= vir_overdens rho_mean volmin ii1^3 whose inner neighbour
//          holds less than the total, 0 if none (tm_ii1 = tm[max(fb, 1)]).
// A non-finite ratio (from non-finite input) takes the reference's serial scan over the
// tables tm (cumulative mass) and cn (ratios) instead (serial = true).
__device__ void virial_finish(double *o, const HaloState &S, const HaloParamsDev &P, int ttab,
                              double ep, double mtot, int i1r, int brk, double tm_brk, int fb,
                              double tm_ii1, bool serial, const double *tm, const double *cn)
{
    const double amax = S.amax, bmax = S.bmax, cmax = S.cmax;
    double ek = o[HP_EK];
    double et = ek + ep;
    o[HP_EP] = ep;
    o[HP_ET] = et;
    double mvir = mtot, avir, bvir, cvir, rvir;
    int i1 = max(i1r, 1);
    avir = (double)i1 / (ttab - 1) * amax;
    bvir = (double)i1 / (ttab - 1) * bmax;
    cvir = (double)i1 / (ttab - 1) * cmax;
    rvir = pow(avir * bvir * cvir, 1.0 / 3.0);
    bool overdensity;
    if (!serial) {
        overdensity = true;
        if (brk >= 0) {
            mvir = tm_brk;
            avir = fmin(avir, (double)brk / (ttab - 1) * amax);
            bvir = fmin(bvir, (double)brk / (ttab - 1) * bmax);
            cvir = fmin(cvir, (double)brk / (ttab - 1) * cmax);
            rvir = pow(avir * bvir * cvir, 1.0 / 3.0);
            overdensity = mvir < P.vir_overdens * P.rho_mean * 4. / 3. * PI * avir * bvir * cvir;
        }
    } else {
        double virth_old = 1.0, virth = 100.0;
        for (int i0 = ttab - 1; i0 >= 0; --i0) {
            if (cn[i0] == -1.0) continue;
            virth = cn[i0];
            if (virth < virth_old) {
                mvir = tm[i0];
                avir = fmin(avir, (double)i0 / (ttab - 1) * amax);
                bvir = fmin(bvir, (double)i0 / (ttab - 1) * bmax);
                cvir = fmin(cvir, (double)i0 / (ttab - 1) * cmax);
                rvir = pow(avir * bvir * cvir, 1.0 / 3.0);
                virth_old = virth;
                if (virth <= 0.20) break;
            }
        }
        overdensity = (virth > 0.20) ||
                      (mvir < P.vir_overdens * P.rho_mean * 4. / 3. * PI * avir * bvir * cvir);
    }
    if (overdensity) {
        const int ii1 = max(fb, 1);
        mvir = tm_ii1;
        avir = (double)ii1 / (ttab - 1) * amax;
        bvir = (double)ii1 / (ttab - 1) * bmax;
        cvir = (double)ii1 / (ttab - 1) * cmax;
        rvir = pow(avir * bvir * cvir, 1.0 / 3.0);
    }
    if (mvir > 0.0 && rvir > 0.0) {
        double rv = fmin(rvir, o[HP_R]);
        o[HP_RVIR] = rv;
        o[HP_MVIR] = mvir;
        o[HP_CVEL] = sqrt(P.gravconst * mvir / rv);
        o[HP_TVIR] = 35.9 * P.gravconst * mvir / rv;
        // the truncated isothermal sphere profile of compute_halo_profile_1b41 (:1232-1238),
        // defined at rvir
        o[HP_RHO0] = mvir / (4.0 * PI * pow(rv, 3.0));
        o[HP_RC] = rv;
    } else {
        o[HP_RVIR] = o[HP_MVIR] = o[HP_CVEL] = o[HP_TVIR] = o[HP_RHO0] = o[HP_RC] = NAN;
    }
    double hl = sqrt(o[HP_LX] * o[HP_LX] + o[HP_LY] * o[HP_LY] + o[HP_LZ] * o[HP_LZ]);
    o[HP_SPIN] = hl * sqrt(fabs(et)) / pow(S.m, 2.5) / P.gravconst;
}

// The shell tables' running sums are sequential (np.cumsum's order, :1024-1025, and the
// potential profile's cumsum, :1090-1103): a chain of nr dependent FP64 additions per table.
// One halo per lane runs its chains side by side with 63 other halos' (a workgroup per halo,
// where one thread added while the others waited, took 2.3 ms of chains at the 50 000-halo
// workload).  The per-tile histograms of a halo of several tiles are first folded, in tile
// order, into the row of its first tile (k_hist_fold), so that a lane reads one row.
template <int BS>
__global__ __launch_bounds__(BS) void k_hist_fold(const int32_t *halos, const int32_t *tile_first,
                                                  int nr, double *hist)
{
    // grid: (halo, chunk of BS entries of its 3 nr); the tiles' values loaded 8 at a time,
    // added in tile order
    const int h = halos[blockIdx.x];
    const int b = blockIdx.y * BS + threadIdx.x;
    if (b >= 3 * nr) return;
    const int t0 = tile_first[h], t1 = tile_first[h + 1];
    const int64_t ld = 3 * (int64_t)nr;
    double *row = hist + t0 * ld;
    double v = row[b];
    int t = t0 + 1;
    for (; t + 8 <= t1; t += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = hist[(t + u) * ld + b];
#pragma unroll
        for (int u = 0; u < 8; ++u) v += x[u];
    }
    for (; t < t1; ++t) v += hist[t * ld + b];
    row[b] = v;
}

// Potential-profile term of shell b (potential_profile_1b40 :1090-1103): dm from the cumulative
// table (np.diff with prepend 0), mass before the shell = cumulative - dm.
__device__ __forceinline__ double shell_scale(int b, int nr)
{
    return (b == 0) ? 0.5 / (double)(nr - 1) : (double)b / (double)(nr - 1);
}

// self-interaction fraction of a shell of cnt members ('count'; 'half_shell' takes 0.5)
__device__ __forceinline__ double shell_self(const HaloParamsDev &P, double cnt)
{
    return P.shell_half ? 0.5 : (cnt > 0 ? 0.5 * (cnt - 1.0) / cnt : 0.0);
}

__device__ __forceinline__ double shell_pe_term(double cum, double prev, double sf, double scale,
                                                double reff0)
{
    const double dm = cum - prev;
    const double before = cum - dm;
    return (before + sf * dm) * dm / (scale * reff0);
}

// ---- shell tables, potential profile and virial scan, one halo per lane.  Pass 1 runs the
// running sums to their ends (the total mass and the profile's outer value, which scale the
// profile and bound the scans); pass 2 runs the same sums again, shell by shell, for each
// shell's virialisation ratio |2K+W|/|K+W| (-1 for an unbound shell, K+W >= 0) and the scans.
// A halo whose ratios hold a NaN is flagged for k_virial_serial.
constexpr int VIR_CH = 8;      // shells loaded at a time per table
constexpr int VIR_HPW = 64;    // halos per wave (one per lane)
constexpr int VIR_SFT = 1024;  // shells of fewer members take their self fraction from a table
template <int BS>
__global__ __launch_bounds__(BS) void k_virial_lanes(int nh, const int32_t *tile_first,
                                                     const int64_t *off, const double *hist,
                                                     int nr, const HaloState *st,
                                                     const double *ep_direct, HaloParamsDev P,
                                                     double *out, int32_t *serial)
{
    // the shells' scales r_b / r_max, the same for every halo, and the self-interaction
    // fractions of shells of fewer than VIR_SFT members (each value the same expression)
    extern __shared__ double sc_[];   // nr + VIR_SFT doubles
    double *const sft_ = sc_ + nr;
    for (int b = threadIdx.x; b < nr; b += BS) sc_[b] = shell_scale(b, nr);
    for (int c = threadIdx.x; c < VIR_SFT; c += BS) sft_[c] = shell_self(P, (double)c);
    __syncthreads();
    const int h = blockIdx.x * VIR_HPW + threadIdx.x;
    if (threadIdx.x >= VIR_HPW || h >= nh) return;
    const double *hm = hist + (int64_t)tile_first[h] * 3 * nr;
    const double *hk = hm + nr;
    const double *hc = hk + nr;
    const HaloState &S = st[h];
    double *o = out + (int64_t)h * HP_NFIELDS;
    const double reff0 = pow(S.amax * S.bmax * S.cmax, 1.0 / 3.0);
    // RF of the axes (det_halo_energies_1b3 :1409-1412 and the 'baseline' profile :1085-1087)
    const double rf = (P.pe_baseline || P.pe_norm_legacy)
                          ? carlson_rf(o[HP_SHA] * o[HP_SHA], o[HP_SHB] * o[HP_SHB],
                                       o[HP_SHC] * o[HP_SHC])
                          : 0.0;
    // one step of the three running sums: cumulative mass cm, kinetic energy ck, and the
    // profile: -3/5 G M(<r)^2 RF itself for 'baseline' (no sum), else the cumulative sum cp of
    // the shell terms
#define VIR_STEP(u, b)                                                                         \
    do {                                                                                       \
        prev = cm;                                                                             \
        cm += m_[u];                                                                           \
        ck += k_[u];                                                                           \
        if (P.pe_baseline) {                                                                   \
            cp = -3.0 / 5.0 * P.gravconst * (cm * cm) * rf;                                    \
        } else {                                                                               \
            const double cnt_ = c_[u];                                                         \
            const double sf_ = cnt_ < (double)VIR_SFT ? sft_[(int)cnt_] : shell_self(P, cnt_);  \
            cp += shell_pe_term(cm, prev, sf_, sc_[b], reff0);                                 \
        }                                                                                      \
    } while (0)
    // a chunk of VIR_CH shells per table (16-byte loads: rows and chunks are 16-byte aligned);
    // the last, partial chunk element by element
#define VIR_LOAD(b0)                                                                           \
    do {                                                                                       \
        if ((b0) + VIR_CH <= nr) {                                                             \
            _Pragma("unroll") for (int u = 0; u < VIR_CH; u += 2) {                            \
                const double2 mm_ = *reinterpret_cast<const double2 *>(hm + (b0) + u);         \
                const double2 kk_ = *reinterpret_cast<const double2 *>(hk + (b0) + u);         \
                const double2 cc_ = *reinterpret_cast<const double2 *>(hc + (b0) + u);         \
                m_[u] = mm_.x;                                                                 \
                m_[u + 1] = mm_.y;                                                             \
                k_[u] = kk_.x;                                                                 \
                k_[u + 1] = kk_.y;                                                             \
                c_[u] = cc_.x;                                                                 \
                c_[u + 1] = cc_.y;                                                             \
            }                                                                                  \
        } else {                                                                               \
            _Pragma("unroll") for (int u = 0; u < VIR_CH; ++u) {                               \
                const int bb_ = min((b0) + u, nr - 1);                                         \
                m_[u] = hm[bb_];                                                               \
                k_[u] = hk[bb_];                                                               \
                c_[u] = hc[bb_];                                                               \
            }                                                                                  \
        }                                                                                      \
    } while (0)
    double m_[VIR_CH], k_[VIR_CH], c_[VIR_CH];
    double cm = 0.0, ck = 0.0, cp = 0.0, prev = 0.0;
    // full chunks without per-shell tests, then the rest
    const int nfull = nr - nr % VIR_CH;
    int b0 = 0;
    for (; b0 < nfull; b0 += VIR_CH) {
        VIR_LOAD(b0);
#pragma unroll
        for (int u = 0; u < VIR_CH; ++u) VIR_STEP(u, b0 + u);
    }
    if (b0 < nr) {
        VIR_LOAD(b0);
#pragma unroll
        for (int u = 0; u < VIR_CH; ++u)
            if (b0 + u < nr) VIR_STEP(u, b0 + u);
    }
    const double mtot = cm;
    const int64_t count = off[h + 1] - off[h];
    // the profile's outer value: -G times the cumulative sum, or the baseline profile's last
    const double gp = P.pe_baseline ? 1.0 : -P.gravconst;
    const double outer = gp * cp;
    const bool small = count < P.direct_pe_nmax;
    // the halo's potential energy: the direct pair sum below the pair-sum limit; above it the
    // profile's outer value ('profile_outer_for_large') or the RF estimate ('legacy')
    const double ep = small ? ep_direct[h]
                            : (P.pe_norm_legacy ? -3.0 / 5.0 * P.gravconst * (S.m * S.m) * rf : outer);
    const bool renorm = (small || P.pe_norm_legacy) && ep != outer && outer != 0.0;
    const int ttab = nr;
    const double volmin = 4. / 3. * PI * (S.amax / (ttab - 1)) * (S.bmax / (ttab - 1)) * (S.cmax / (ttab - 1));
    // vir_overdens * rho_mean * volmin * ii1**3, the cube an exact integer
    const double mv0 = P.vir_overdens * P.rho_mean * volmin;
    int i1r = 0, brk = -1, fb = 0;
    bool nan = false;
    double tm_brk = 0.0, tm_fb = 0.0, tm1 = 0.0;
    cm = ck = cp = prev = 0.0;
    // one shell of pass 2: its ratio and the scans
#define VIR_SCAN(u, b)                                                                         \
    do {                                                                                       \
        VIR_STEP(u, b);                                                                        \
        /* profile (rescaled to the halo's potential energy where the policy says so) */       \
        double pr = gp * cp;                                                                   \
        if (renorm) pr = pr / outer * ep;                                                      \
        const double kw = ck + pr;                                                             \
        const double v = (kw >= 0.0) ? -1.0 : fabs((2.0 * ck + pr) / kw);                     \
        if (v != v) nan = true;                                                                \
        else if (v != -1.0 && v <= 0.20) {                                                     \
            brk = (b);                                                                         \
            tm_brk = cm;                                                                       \
        }                                                                                      \
        if ((b) >= 1) {                                                                        \
            if (prev < cm) i1r = (b);                                                          \
            if (cm >= mv0 * (double)((int64_t)(b) * (b) * (b)) && prev < mtot) {               \
                fb = (b);                                                                      \
                tm_fb = cm;                                                                    \
            }                                                                                  \
        }                                                                                      \
        if ((b) == 1) tm1 = cm;                                                                \
    } while (0)
    b0 = 0;
    for (; b0 < nfull; b0 += VIR_CH) {
        VIR_LOAD(b0);
#pragma unroll
        for (int u = 0; u < VIR_CH; ++u) VIR_SCAN(u, b0 + u);
    }
    if (b0 < nr) {
        VIR_LOAD(b0);
#pragma unroll
        for (int u = 0; u < VIR_CH; ++u)
            if (b0 + u < nr) VIR_SCAN(u, b0 + u);
    }
#undef VIR_SCAN
#undef VIR_LOAD
#undef VIR_STEP
    if (nan) {
        serial[h] = 1;
        return;
    }
    virial_finish(o, S, P, ttab, ep, mtot, i1r, brk, tm_brk, fb, fb >= 1 ? tm_fb : tm1, false,
                  nullptr, nullptr);
}

// ---- the same, one workgroup per halo, for the halos k_virial_lanes flagged (a NaN ratio
// takes the reference's serial scan over the tables): the tables in LDS, the running sums by
// one thread each, the scans as block-wide maximum-index searches.  The other workgroups
// return at once.
template <int BS>
__global__ __launch_bounds__(BS) void k_virial_serial(int nh, const int32_t *tile_first,
                                                      const int64_t *off, const double *hist,
                                                      int nr, const HaloState *st,
                                                      const double *ep_direct, HaloParamsDev P,
                                                      double *out, const int32_t *serial)
{
    extern __shared__ double work[];          // 4*nr doubles
    const int h = blockIdx.x;
    if (h >= nh || !serial[h]) return;
    double *tm = work;                         // shell mass -> cumulative mass
    double *tk = tm + nr;                      // shell kinetic energy -> cumulative
    double *cn = tk + nr;                      // shell counts -> virialisation ratio per shell
    double *tp = cn + nr;                      // potential profile
    const HaloState &S = st[h];
    // the halo's histograms, folded into its first tile's row
    const double *row = hist + (int64_t)tile_first[h] * 3 * nr;
    for (int b = threadIdx.x; b < nr; b += BS) {
        tm[b] = row[b];
        tk[b] = row[nr + b];
        cn[b] = row[2 * nr + b];
    }
    __syncthreads();
    if (threadIdx.x == 0) seq_cumsum(tm, nr);
    else if (threadIdx.x == WAVE_HP) seq_cumsum(tk, nr);
    __syncthreads();
    const double reff0 = pow(S.amax * S.bmax * S.cmax, 1.0 / 3.0);
    double *o = out + (int64_t)h * HP_NFIELDS;
    const double rf = (P.pe_baseline || P.pe_norm_legacy)
                          ? carlson_rf(o[HP_SHA] * o[HP_SHA], o[HP_SHB] * o[HP_SHB],
                                       o[HP_SHC] * o[HP_SHC])
                          : 0.0;
    if (P.pe_baseline) {
        for (int b = threadIdx.x; b < nr; b += BS) tp[b] = -3.0 / 5.0 * P.gravconst * (tm[b] * tm[b]) * rf;
    } else {
        for (int b = threadIdx.x; b < nr; b += BS)
            tp[b] = shell_pe_term(tm[b], b > 0 ? tm[b - 1] : 0.0, shell_self(P, cn[b]),
                                  shell_scale(b, nr), reff0);
        __syncthreads();
        if (threadIdx.x == 0) seq_cumsum(tp, nr);
    }
    __syncthreads();
    const int64_t count = off[h + 1] - off[h];
    const double gp = P.pe_baseline ? 1.0 : -P.gravconst;
    const double outer = gp * tp[nr - 1];
    const bool small = count < P.direct_pe_nmax;
    const double ep = small ? ep_direct[h]
                            : (P.pe_norm_legacy ? -3.0 / 5.0 * P.gravconst * (S.m * S.m) * rf : outer);
    const bool renorm = (small || P.pe_norm_legacy) && ep != outer && outer != 0.0;
    __syncthreads();
    for (int b = threadIdx.x; b < nr; b += BS) {
        double p = gp * tp[b];
        if (renorm) p = p / outer * ep;
        tp[b] = p;
        const double kw = tk[b] + p;
        cn[b] = (kw >= 0.0) ? -1.0 : fabs((2.0 * tk[b] + p) / kw);
    }
    __syncthreads();
    const int ttab = nr;
    const double volmin = 4. / 3. * PI * (S.amax / (ttab - 1)) * (S.bmax / (ttab - 1)) * (S.cmax / (ttab - 1));
    const double mv0 = P.vir_overdens * P.rho_mean * volmin;
    const double mtot = tm[ttab - 1];
    __shared__ int red[4][BS];
    {
        int b_i1 = 0, b_brk = -1, b_fb = 0, b_nan = 0;
        for (int b = threadIdx.x; b < ttab; b += BS) {
            const double v = cn[b];
            if (v != v) b_nan = 1;
            else if (v != -1.0 && v <= 0.20) b_brk = b;
            if (b >= 1) {
                if (tm[b - 1] < tm[b]) b_i1 = b;
                if (tm[b] >= mv0 * (double)((int64_t)b * b * b) && tm[b - 1] < mtot) b_fb = b;
            }
        }
        const int t = threadIdx.x;
        red[0][t] = b_i1;
        red[1][t] = b_brk;
        red[2][t] = b_fb;
        red[3][t] = b_nan;
        __syncthreads();
        for (int w = BS / 2; w > 0; w >>= 1) {
            if (t < w)
                for (int q = 0; q < 4; ++q) red[q][t] = max(red[q][t], red[q][t + w]);
            __syncthreads();
        }
    }
    if (threadIdx.x != 0) return;
    const int brk = red[1][0], fb = red[2][0];
    virial_finish(o, S, P, ttab, ep, mtot, red[0][0], brk, brk >= 0 ? tm[brk] : 0.0, fb,
                  tm[max(fb, 1)], red[3][0] != 0, tm, cn);
}

template <typename T>
T *dalloc(std::vector<void *> &keep, size_t n)
{
    void *p = nullptr;
    AHOP_HIP(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
    keep.push_back(p);
    return (T *)p;
}

struct Tiles {
    std::vector<int32_t> halo, first;
    std::vector<int64_t> start, end;
};

Tiles make_tiles(const std::vector<int64_t> &off, int64_t T)
{
    Tiles t;
    int nh = (int)off.size() - 1;
    t.first.resize(nh + 1);
    for (int h = 0; h < nh; ++h) {
        t.first[h] = (int32_t)t.halo.size();
        for (int64_t s = off[h]; s < off[h + 1]; s += T) {
            t.halo.push_back(h);
            t.start.push_back(s);
            t.end.push_back(std::min(s + T, off[h + 1]));
        }
    }
    t.first[nh] = (int32_t)t.halo.size();
    return t;
}

}  // namespace

// SC on the listed halos through their member tiles (ScState per halo; part / pkey / pidx
// are the per-tile scratch of the member passes, free at this point).  One stream sync: the
// starting radii come back to the host, which counts the spheres as the kernels will.
static void sc_tiled(const std::vector<int32_t> &big, const Tiles &tl, const HaloDev &D,
                     const int32_t *d_tf, const int64_t *d_off, const double *pos, HaloState *st,
                     const HaloParamsDev &P, double *part, double *pkey, int64_t *pidx,
                     std::vector<void *> &keep, std::vector<int32_t> &tiles, int *err,
                     hipStream_t s)
{
    const int nb = (int)big.size();
    for (int32_t h : big)
        for (int32_t t = tl.first[h]; t < tl.first[h + 1]; ++t) tiles.push_back(t);
    const unsigned ntl = (unsigned)tiles.size(), cb = (unsigned)((nb + 63) / 64);
    int32_t *d_big = dalloc<int32_t>(keep, nb), *d_tiles = dalloc<int32_t>(keep, ntl);
    ScState *sc = dalloc<ScState>(keep, (size_t)tl.first.size());
    double *d_r0 = dalloc<double>(keep, nb);
    AHOP_HIP(hipMemcpyAsync(d_big, big.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s));
    AHOP_HIP(hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int32_t) * ntl, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_sc_radius_tiles<256>, dim3(ntl), dim3(256), 0, s, D, st, d_tiles, pkey);
    hipLaunchKernelGGL(k_sc_init, dim3(cb), dim3(64), 0, s, d_big, nb, d_tf, pkey, st, sc, d_r0);
    std::vector<double> r0(nb);
    AHOP_HIP(hipMemcpyAsync(r0.data(), d_r0, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    // spheres while r > dcell_min, then the last one: the same double arithmetic as the kernels
    int nsph = 0;
    for (double r : r0) {
        if (!(std::isfinite(r) && r > 0.0)) continue;
        int n = 0;
        while (r > P.dcell_min && n < RC_MAX_SPHERES) {
            ++n;
            r = (1 - P.eps_sc) * r;
        }
        if (n >= RC_MAX_SPHERES) {
            const int flag = RC_TOO_DEEP;
            AHOP_HIP(hipMemcpyAsync(err, &flag, sizeof(int), hipMemcpyHostToDevice, s));
            AHOP_HIP(hipStreamSynchronize(s));
            return;
        }
        nsph = std::max(nsph, n);
    }
    for (int it = 0; it <= nsph; ++it) {
        hipLaunchKernelGGL(k_sc_sphere_tiles<256>, dim3(ntl), dim3(256), 0, s, D, sc, d_tiles, part);
        hipLaunchKernelGGL(k_sc_sphere_combine, dim3(cb), dim3(64), 0, s, d_big, nb, d_tf, part, sc,
                           P.dcell_min, P.eps_sc);
    }
    hipLaunchKernelGGL(k_sc_nearest_tiles<256>, dim3(ntl), dim3(256), 0, s, D, sc, d_tiles, pkey, pidx);
    hipLaunchKernelGGL(k_sc_nearest_combine, dim3(cb), dim3(64), 0, s, d_big, nb, d_tf, pkey, pidx,
                       d_off, pos, sc, st, err);
}

// Runs every pass for one tile size / workgroup size; returns device milliseconds, and in
// sub_ms (may be null) those of the pair-sum potential (k_direct_pe) and of the shell tables
// and virial scan (k_hist_fold, k_virial_lanes, k_virial_serial).
template <int BS>
static float run_passes(const std::vector<int64_t> &hoff, const int64_t *d_off, const double *pos,
                        const double *vel, const double *mass, const double *dens, int64_t T,
                        const HaloParamsDev &P, double *d_out, hipStream_t s, bool full,
                        float *sub_ms = nullptr)
{
    const int nh = (int)hoff.size() - 1;
    Tiles tl = make_tiles(hoff, T);
    const int nt = (int)tl.halo.size();
    std::vector<int32_t> small;   // halos of the pair-sum branch (host copy kept until the sync)
    std::vector<int32_t> multi;   // halos of several tiles (likewise)
    std::vector<int32_t> cls[4];  // recentring size classes (likewise)
    std::vector<int32_t> sc_tiles;
    std::vector<void *> keep;
    auto cleanup = [&]() {
        for (void *p : keep) (void)hipFree(p);
        keep.clear();
    };
    try {
        int32_t *d_th = dalloc<int32_t>(keep, nt), *d_tf = dalloc<int32_t>(keep, nh + 1);
        int64_t *d_ts = dalloc<int64_t>(keep, nt), *d_te = dalloc<int64_t>(keep, nt);
        AHOP_HIP(hipMemcpyAsync(d_th, tl.halo.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_tf, tl.first.data(), sizeof(int32_t) * (nh + 1), hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_ts, tl.start.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice, s));
        AHOP_HIP(hipMemcpyAsync(d_te, tl.end.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice, s));
        HaloDev D{d_off, pos, vel, mass, dens, d_th, d_ts, d_te, P.lbox, P.hub_pt, P.gravconst};
        double *part = dalloc<double>(keep, (size_t)nt * NSUM3);
        double *pkey = dalloc<double>(keep, nt);
        int64_t *pidx = dalloc<int64_t>(keep, nt);
        HaloState *st = dalloc<HaloState>(keep, nh);
        int *missed = dalloc<int>(keep, 2);            // [0]: shell misses, [1]: recentring errors
        AHOP_HIP(hipMemsetAsync(missed, 0, 2 * sizeof(int), s));
        const unsigned hb = (unsigned)((nh + 127) / 128);
        hipEvent_t e0, e1, ep0, ep1;
        AHOP_HIP(hipEventCreate(&e0));
        AHOP_HIP(hipEventCreate(&e1));
        AHOP_HIP(hipEventCreate(&ep0));
        AHOP_HIP(hipEventCreate(&ep1));
        bool timed_pe = false;
        AHOP_HIP(hipEventRecord(e0, s));
        if (P.light)   // the columns of the skipped stages read NaN (all-ones bytes)
            AHOP_HIP(hipMemsetAsync(d_out, 0xFF, sizeof(double) * HP_NFIELDS * nh, s));
        hipLaunchKernelGGL(k_pass1<BS>, dim3(nt), dim3(BS), 0, s, D, part);
        hipLaunchKernelGGL(k_combine1, dim3(hb), dim3(128), 0, s, nh, d_tf, part, d_off, pos, P.lbox, st);
        hipLaunchKernelGGL(k_pass2<BS>, dim3(nt), dim3(BS), 0, s, D, st, P.cdm, pkey, pidx);
        hipLaunchKernelGGL(k_combine2, dim3(hb), dim3(128), 0, s, nh, d_tf, pkey, pidx, d_off, pos, st);
        if (full && (P.sc || P.dpmmc)) {
            // SC alone: halos from SC_TILED members on take the tiled sphere path (cls[3]);
            // the others one workgroup per halo, sized to the halo: one wave below 2048
            // members, 4 waves below 32768, 16 above (every sphere is a pass over the halo)
            const bool tiled = P.sc && !P.dpmmc;
            for (int h = 0; h < nh; ++h) {
                const int64_t c = hoff[h + 1] - hoff[h];
                cls[(tiled && c >= SC_TILED) ? 3 : c < 2048 ? 0 : c < 32768 ? 1 : 2].push_back(h);
            }
            if (!cls[3].empty())
                sc_tiled(cls[3], tl, D, d_tf, d_off, pos, st, P, part, pkey, pidx, keep, sc_tiles,
                         missed + 1, s);
            int32_t *d_cls = dalloc<int32_t>(keep, nh);
            const size_t glds = P.dpmmc ? sizeof(double) * 729 : 0;
            int pos0 = 0;
            for (int k = 2; k >= 0; --k) {             // largest first: they set the tail
                const unsigned nk = (unsigned)cls[k].size();
                if (!nk) continue;
                AHOP_HIP(hipMemcpyAsync(d_cls + pos0, cls[k].data(), sizeof(int32_t) * nk,
                                        hipMemcpyHostToDevice, s));
                if (k == 2)
                    hipLaunchKernelGGL(k_recenter<1024>, dim3(nk), dim3(1024), glds, s, D, st,
                                       d_cls + pos0, P.sc, P.dpmmc, P.dcell_min, P.eps_sc, missed + 1);
                else if (k == 1)
                    hipLaunchKernelGGL(k_recenter<256>, dim3(nk), dim3(256), glds, s, D, st,
                                       d_cls + pos0, P.sc, P.dpmmc, P.dcell_min, P.eps_sc, missed + 1);
                else
                    hipLaunchKernelGGL(k_recenter<64>, dim3(nk), dim3(64), glds, s, D, st,
                                       d_cls + pos0, P.sc, P.dpmmc, P.dcell_min, P.eps_sc, missed + 1);
                pos0 += (int)nk;
            }
        }
        hipLaunchKernelGGL(k_pass3<BS>, dim3(nt), dim3(BS), 0, s, D, st, part);
        hipLaunchKernelGGL(k_combine3, dim3(hb), dim3(128), 0, s, nh, d_tf, part, P.gravconst, st, d_out);
        if (full && !P.light) {
            hipLaunchKernelGGL(k_pass4a<BS>, dim3(nt), dim3(BS), 0, s, D, st, pkey);
            hipLaunchKernelGGL(k_combine4a, dim3(hb), dim3(128), 0, s, nh, d_tf, pkey, st);
            const int nr = P.nshell;
            double *hist = dalloc<double>(keep, (size_t)nt * 3 * nr);
            hipLaunchKernelGGL(k_pass4b<BS>, dim3(nt), dim3(BS), sizeof(double) * 3 * nr, s, D, st,
                               nr, hist, missed);
            int64_t nsmall_max = 0;
            for (int h = 0; h < nh; ++h)
                if (hoff[h + 1] - hoff[h] < P.direct_pe_nmax) {
                    small.push_back(h);
                    nsmall_max = std::max(nsmall_max, hoff[h + 1] - hoff[h]);
                }
            double *ep = dalloc<double>(keep, nh);
            int32_t *d_small = nullptr;
            if (!small.empty()) {
                if (nsmall_max > PE_NMAX) throw std::runtime_error("direct_pe_nmax above the LDS limit");
                d_small = dalloc<int32_t>(keep, small.size());
                AHOP_HIP(hipMemcpyAsync(d_small, small.data(), sizeof(int32_t) * small.size(),
                                        hipMemcpyHostToDevice, s));
            }
            AHOP_HIP(hipEventRecord(ep0, s));
            if (!small.empty()) {
                hipLaunchKernelGGL((k_direct_pe<256, false>), dim3((unsigned)small.size()),
                                   dim3(256), sizeof(double4) * nsmall_max, s, D, d_small, ep);
                hipLaunchKernelGGL((k_direct_pe<256, true>), dim3((unsigned)small.size()),
                                   dim3(256), sizeof(double4) * nsmall_max, s, D, d_small, ep);
            }
            AHOP_HIP(hipEventRecord(ep1, s));
            timed_pe = true;
            for (int h = 0; h < nh; ++h)
                if (tl.first[h + 1] - tl.first[h] > 1) multi.push_back(h);
            int32_t *d_multi = nullptr;
            if (!multi.empty()) {
                d_multi = dalloc<int32_t>(keep, multi.size());
                AHOP_HIP(hipMemcpyAsync(d_multi, multi.data(), sizeof(int32_t) * multi.size(),
                                        hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_hist_fold<256>,
                                   dim3((unsigned)multi.size(), (unsigned)((3 * nr + 255) / 256)),
                                   dim3(256), 0, s, d_multi, d_tf, nr, hist);
            }
            int32_t *d_serial = dalloc<int32_t>(keep, nh);
            AHOP_HIP(hipMemsetAsync(d_serial, 0, sizeof(int32_t) * nh, s));
            hipLaunchKernelGGL(k_virial_lanes<64>, dim3((nh + VIR_HPW - 1) / VIR_HPW), dim3(64),
                               sizeof(double) * (nr + VIR_SFT), s, nh, d_tf, d_off, hist, nr, st,
                               ep, P, d_out, d_serial);
            hipLaunchKernelGGL(k_virial_serial<256>, dim3(nh), dim3(256), sizeof(double) * 4 * nr, s,
                               nh, d_tf, d_off, hist, nr, st, ep, P, d_out, d_serial);
            if (P.angmom_nshells > 0) {
                const int nsh = P.angmom_nshells;
                double *ahist = dalloc<double>(keep, (size_t)nt * 4 * nsh);
                hipLaunchKernelGGL(k_angmom<BS>, dim3(nt), dim3(BS), sizeof(double) * 4 * nsh, s,
                                   D, st, nsh, ahist, missed);
                hipLaunchKernelGGL(k_angmom_combine, dim3(nh), dim3(128), 0, s, nh, d_tf, ahist,
                                   nsh, st, d_out, P.angmom, P.angmom_axes);
            }
        }
        AHOP_HIP(hipEventRecord(e1, s));
        AHOP_HIP(hipStreamSynchronize(s));
        float ms = 0;
        AHOP_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (sub_ms) {
            sub_ms[0] = sub_ms[1] = 0.0f;
            if (timed_pe) {
                AHOP_HIP(hipEventElapsedTime(&sub_ms[0], ep0, ep1));
                AHOP_HIP(hipEventElapsedTime(&sub_ms[1], ep1, e1));
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(ep0);
        (void)hipEventDestroy(ep1);
        int hmissed[2] = {0, 0};
        AHOP_HIP(hipMemcpy(hmissed, missed, 2 * sizeof(int), hipMemcpyDeviceToHost));
        cleanup();
        if (hmissed[1] & RC_DPMMC_INDEX)
            throw std::runtime_error("DPMMC centre: the densest member's index is beyond the "
                                     "members inside the last grid (the reference's IndexError)");
        if (hmissed[1] & RC_SC_DISTMIN)
            throw std::runtime_error("SC centre: squared distance to the nearest member above 1 "
                                     "(the reference's ValueError)");
        if (hmissed[1] & RC_TOO_DEEP)
            throw std::runtime_error("recentring: more than " + std::to_string(RC_MAX_LEVELS) +
                                     " grid levels or " + std::to_string(RC_MAX_SPHERES) +
                                     " spheres (raise dcell_min or eps_SC)");
        if (hmissed[0])
            throw std::runtime_error("tab_props_inside: " + std::to_string(hmissed[0]) +
                                     " members fall outside the outer ellipsoid");
        return ms;
    } catch (...) {
        cleanup();
        throw;
    }
}

HaloTuning halo_props_device(const int64_t *h_off, int64_t nhalos, const double *pos,
                             const double *vel, const double *mass, const double *dens,
                             const HaloParamsDev &P, double *d_out, int tile, int bs,
                             hipStream_t s)
{
    if (nhalos <= 0) throw std::runtime_error("need at least one halo");
    // the shell tables of one halo live in LDS (4 doubles per shell)
    if (P.nshell < 2 || P.nshell > 2000) throw std::runtime_error("nshell must be in [2, 2000]");
    if (P.angmom_nshells < 0 || P.angmom_nshells > 2000)
        throw std::runtime_error("the per-shell angular momentum needs 1 to 2000 shells");
    if (P.angmom_nshells > 0 && (P.light || !P.angmom || !P.angmom_axes))
        throw std::runtime_error("the per-shell angular momentum needs the full mode and its "
                                 "two output arrays");
    if (!P.cdm && !dens) throw std::runtime_error("densest-member centre needs densities");
    if (P.dpmmc && !dens) throw std::runtime_error("DPMMC recentring needs densities");
    if ((P.sc || P.dpmmc) && !(P.dcell_min > 0))
        throw std::runtime_error("SC and DPMMC recentring need dcell_min > 0");
    if (P.sc && !(P.eps_sc > 0 && P.eps_sc < 1))
        throw std::runtime_error("SC recentring needs 0 < eps_SC < 1");
    std::vector<int64_t> hoff(h_off, h_off + nhalos + 1);
    if (hoff[0] < 0) throw std::runtime_error("negative member offset");
    if (nhalos > INT32_MAX / 2) throw std::runtime_error("too many halos");
    for (int64_t h = 0; h < nhalos; ++h)
        if (hoff[h + 1] <= hoff[h]) throw std::runtime_error("every halo needs at least one member");
    const int64_t M = hoff[nhalos] - hoff[0];
    int64_t *d_off = nullptr;
    AHOP_HIP(hipMalloc(&d_off, sizeof(int64_t) * (nhalos + 1)));
    AHOP_HIP(hipMemcpyAsync(d_off, hoff.data(), sizeof(int64_t) * (nhalos + 1), hipMemcpyHostToDevice, s));
    HaloTuning tune{};
    try {
        if (tile <= 0 || bs <= 0) {
            // autotune the member-pass schedule on this workload: time passes 1-3 for each
            // candidate (tile, workgroup) and keep the fastest
            const int64_t tiles[] = {1024, 4096, 16384, 65536};
            const int wgs[] = {128, 256};
            float best = 1e30f;
            for (int64_t T : tiles)
                for (int w : wgs) {
                    if (T > std::max<int64_t>(M, 1024) * 4) continue;
                    float ms = (w == 128) ? run_passes<128>(hoff, d_off, pos, vel, mass, dens, T, P, d_out, s, false)
                                          : run_passes<256>(hoff, d_off, pos, vel, mass, dens, T, P, d_out, s, false);
                    tune.trials++;
                    if (ms < best) {
                        best = ms;
                        tune.tile = (int)T;
                        tune.bs = w;
                    }
                }
        } else {
            tune.tile = tile;
            tune.bs = bs;
        }
        if (tune.bs == 128)
            tune.ms = run_passes<128>(hoff, d_off, pos, vel, mass, dens, tune.tile, P, d_out, s, true,
                                      tune.sub_ms);
        else if (tune.bs == 256)
            tune.ms = run_passes<256>(hoff, d_off, pos, vel, mass, dens, tune.tile, P, d_out, s, true,
                                      tune.sub_ms);
        else
            throw std::runtime_error("workgroup size must be 128 or 256");
    } catch (...) {
        (void)hipFree(d_off);
        throw;
    }
    (void)hipFree(d_off);
    return tune;
}

}  // namespace ahop
