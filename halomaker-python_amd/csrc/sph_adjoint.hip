// sph_adjoint.hip -- hand-written adjoint of the SPH density with frozen neighbour lists.
//
// Forward (omp_compute_density, fortran/compute_adaptahop.f90:2816-2838, spline :3270-3287),
// with the neighbour rows j_1..j_k of particle i in ascending distance:
//   r_i   = sqrt(d2(i, j_k)) / 2
//   rho_i = A * (sum_{n<k} m_{j_n} W(d_{i j_n} / r_i) + m_i) / r_i^3,   A = L^3 / pi
// Adjoint, for an upstream gradient g_i = dLoss/drho_i:
//   dLoss/dm_i     += g_i A / r_i^3                               (own mass)
//   dLoss/dm_j     += g_i A W(q_ij) / r_i^3                       (each of the first k-1)
//   dLoss/dx_j     += g_i A m_j W'(q_ij) / r_i^4 * u_ij,  dLoss/dx_i -= the same
//   dLoss/dx_{j_k} += g_i (drho_i/dr_i) / 2 * u_ik,       dLoss/dx_i -= the same
//   drho_i/dr_i     = -A / r_i^4 * (3 (C_i + m_i) + S_i / r_i),
//   C_i = sum m_j W(q_ij),  S_i = sum m_j W'(q_ij) d_ij
// where u_ij is the unit minimum-image separation x_j - x_i.  Per-pair terms and sums are
// FP64: with FP32 pair terms the gradient of a mass that many particles list as a neighbour
// loses up to ~3e-4 relative to cancellation (the upstream weights g_i have both signs).
// Each particle gathers its own terms and adds its neighbours' terms with FP64 atomics.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

#include "ahop_common.h"

namespace ahop {

namespace {

__device__ __forceinline__ double min_image(double d, double L)
{
    // the minimum-image separation, the signed counterpart of the reference's
    // min(|d|, L - |d|) (walk_tree2 :2931-2941)
    const double h = 0.5 * L;
    if (d > h) return d - L;
    if (d < -h) return d + L;
    return d;
}

// W(q) and W'(q) of the cubic spline (spline :3270-3287)
__device__ __forceinline__ void spline_wd(double q, double &w, double &wp)
{
    if (q <= 1.0) {
        w = 1.0 - 1.5 * q * q + 0.75 * q * q * q;
        wp = -3.0 * q + 2.25 * q * q;
    } else if (q <= 2.0) {
        const double t = 2.0 - q;
        w = 0.25 * t * t * t;
        wp = -0.75 * t * t;
    } else {
        w = 0.0;
        wp = 0.0;
    }
}

// With an order (WINDOW): workgroup b takes positions [b*ADJ_BS, (b+1)*ADJ_BS) of a
// permutation that keeps neighbours close (the forward's Morton order), and the terms of
// neighbours whose position lies within ADJ_HALF of the workgroup's range are first added into
// LDS (FP64 LDS atomics) and then added to global memory once per position; only the others
// take a global atomic per term.  Without an order, thread i takes particle i and every
// neighbour term is a global atomic.
constexpr int ADJ_BS = 256;
constexpr int ADJ_HALF = 256;
constexpr int ADJ_WIN = ADJ_BS + 2 * ADJ_HALF;   // 768 positions: 24 KiB of (x, y, z, m) sums

template <bool WINDOW>
__global__ __launch_bounds__(ADJ_BS) void k_sph_adjoint(
    const double *__restrict__ pos, const double *__restrict__ mass,
    const int32_t *__restrict__ nbr, const int32_t *__restrict__ order,
    const int32_t *__restrict__ slot_of, const double *__restrict__ grad_rho, int64_t n, int K,
    double Lx, double Ly, double Lz, double A, double *__restrict__ gpos,
    double *__restrict__ gmass)
{
    __shared__ double acc[WINDOW ? 4 * ADJ_WIN : 1];   // [component][window position]
    const int64_t s = blockIdx.x * (int64_t)ADJ_BS + threadIdx.x;
    const int64_t w0 = blockIdx.x * (int64_t)ADJ_BS - ADJ_HALF;
    if constexpr (WINDOW) {
        for (int u = threadIdx.x; u < 4 * ADJ_WIN; u += ADJ_BS) acc[u] = 0.0;
        __syncthreads();
    }
    // add (fx, fy, fz) to the position gradient and fm to the mass gradient of particle j;
    // uj is j's window position (meaningful with WINDOW only)
    auto add = [&](int64_t j, int64_t uj, bool with_pos, double fx, double fy, double fz,
                   double fm) {
        if (WINDOW && uj >= 0 && uj < ADJ_WIN) {
            atomicAdd(&acc[3 * ADJ_WIN + uj], fm);
            if (with_pos) {
                atomicAdd(&acc[uj], fx);
                atomicAdd(&acc[ADJ_WIN + uj], fy);
                atomicAdd(&acc[2 * ADJ_WIN + uj], fz);
            }
        } else {
            unsafeAtomicAdd(&gmass[j], fm);
            if (with_pos) {
                unsafeAtomicAdd(&gpos[3 * j], fx);
                unsafeAtomicAdd(&gpos[3 * j + 1], fy);
                unsafeAtomicAdd(&gpos[3 * j + 2], fz);
            }
        }
    };
    if (s < n) {
        const int64_t i = WINDOW ? (int64_t)order[s] : s;
        const double g = grad_rho[i];
        const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
        const int32_t *row = nbr + i * (int64_t)K;
        // smoothing length from the k-th neighbour, in the reference's FP64 arithmetic
        const int32_t jk = row[K - 1];
        const double kx = min_image(pos[3 * (int64_t)jk] - xi, Lx);
        const double ky = min_image(pos[3 * (int64_t)jk + 1] - yi, Ly);
        const double kz = min_image(pos[3 * (int64_t)jk + 2] - zi, Lz);
        const double dk = sqrt((kx * kx + ky * ky) + kz * kz);
        const double r = 0.5 * dk;
        const double ir = 1.0 / r;
        const double ir3 = ir * ir * ir;
        const double gA3 = g * A * ir3;          // g A / r^3
        const double gA4 = gA3 * ir;             // g A / r^4 (pair-force scale)
        double C = 0.0, S = 0.0, ox = 0.0, oy = 0.0, oz = 0.0;
        for (int t = 0; t < K - 1; ++t) {
            const int64_t j = row[t];
            const int64_t uj = WINDOW ? (int64_t)slot_of[j] - w0 : -1;
            const double dx = min_image(pos[3 * j] - xi, Lx);
            const double dy = min_image(pos[3 * j + 1] - yi, Ly);
            const double dz = min_image(pos[3 * j + 2] - zi, Lz);
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            double w, wp;
            spline_wd(d * ir, w, wp);
            const double mj = mass[j];
            C += mj * w;
            S += mj * wp * d;
            double fx = 0.0, fy = 0.0, fz = 0.0;
            if (d > 0.0) {
                const double c = gA4 * mj * wp / d;   // along x_j - x_i, per unit length
                fx = c * dx;
                fy = c * dy;
                fz = c * dz;
                ox -= fx;
                oy -= fy;
                oz -= fz;
            }
            add(j, uj, d > 0.0, fx, fy, fz, gA3 * w);
        }
        const double mi = mass[i];
        // the k-th neighbour moves the smoothing length: d r / d x_{j_k} = u_ik / 2
        if (dk > 0.0) {
            const double drho_dr = -A * ir3 * ir * (3.0 * (C + mi) + S * ir);
            const double c = g * drho_dr * 0.5 / dk;
            const double fx = c * kx, fy = c * ky, fz = c * kz;
            ox -= fx;
            oy -= fy;
            oz -= fz;
            add(jk, WINDOW ? (int64_t)slot_of[jk] - w0 : -1, true, fx, fy, fz, 0.0);
        }
        add(i, s - w0, true, ox, oy, oz, gA3);   // own terms (i's window position is s - w0)
    }
    if constexpr (WINDOW) {
        __syncthreads();
        for (int u = threadIdx.x; u < ADJ_WIN; u += ADJ_BS) {
            const int64_t sj = w0 + u;
            if (sj < 0 || sj >= n) continue;
            const int64_t j = order[sj];
            const double fm = acc[3 * ADJ_WIN + u], fx = acc[u], fy = acc[ADJ_WIN + u],
                         fz = acc[2 * ADJ_WIN + u];
            if (fm != 0.0) unsafeAtomicAdd(&gmass[j], fm);
            if (fx != 0.0) unsafeAtomicAdd(&gpos[3 * j], fx);
            if (fy != 0.0) unsafeAtomicAdd(&gpos[3 * j + 1], fy);
            if (fz != 0.0) unsafeAtomicAdd(&gpos[3 * j + 2], fz);
        }
    }
}

// slot_of[order[s]] = s (slot_of starts at -1, so entries left at -1 reveal a non-permutation)
__global__ void k_inverse_order(const int32_t *__restrict__ order, int64_t n,
                                int32_t *__restrict__ slot_of)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < n) slot_of[order[s]] = (int32_t)s;
}

__global__ void k_fill_m1(int32_t *__restrict__ a, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = -1;
}

__global__ void k_count_negative(const int32_t *__restrict__ a, int64_t n, int *__restrict__ bad)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && a[i] < 0) atomicAdd(bad, 1);
}

__global__ void k_zero(double *__restrict__ a, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// counts neighbour ids outside [0, n) (the adjoint indexes positions with them)
__global__ void k_count_bad_ids(const int32_t *__restrict__ nbr, int64_t m, int64_t n,
                                int *__restrict__ bad)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < m && (nbr[e] < 0 || nbr[e] >= n)) atomicAdd(bad, 1);
}

}  // namespace

// single-rounded pi of the reference's density (f90:58)
static constexpr double PI_REF_HOST = 3.1415927410125732421875;

namespace {

__global__ void k_set_int(int *__restrict__ a, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *a = v;
}

// Checks and launch of the adjoint with the caller's scratch: bad (one int) and, with an
// order, slot (n int32, the order's inverse).  The neighbour ids must lie in [0, n) and the
// order (if any) be a permutation of [0, n): both are counted on the device and read back
// before the kernels that index with them run.  e0 / e1 (may be null) bracket the kernels.
void check_args(int64_t n, int K, double Lx, double Ly, double Lz)
{
    if (n < 2) throw std::runtime_error("need at least 2 particles");
    if (K < 1 || K >= n) throw std::runtime_error("need 1 <= k < npart");
    if (!(Lx > 0 && Ly > 0 && Lz > 0)) throw std::runtime_error("box sides must be > 0");
    if (n > INT32_MAX) throw std::runtime_error("npart must fit in int32");
}

void adjoint_launch(const double *d_pos, const double *d_mass, const int32_t *d_nbr,
                    const int32_t *d_order, const double *d_grad_rho, int64_t n, int K,
                    double Lx, double Ly, double Lz, double *d_gpos, double *d_gmass, int *d_bad,
                    int32_t *d_slot, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    check_args(n, K, Lx, Ly, Lz);
    const double A = Lx * Ly * Lz / PI_REF_HOST;
    const unsigned b1 = (unsigned)((n + 255) / 256);
    int bad[2] = {0, 0};
    const int64_t m = n * (int64_t)K;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, d_bad, 0);
    hipLaunchKernelGGL(k_count_bad_ids, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, d_nbr,
                       m, n, d_bad);
    AHOP_HIP(hipMemcpyAsync(&bad[0], d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    if (d_order) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, d_bad, 0);
        hipLaunchKernelGGL(k_count_bad_ids, dim3(b1), dim3(256), 0, s, d_order, n, n, d_bad);
        AHOP_HIP(hipMemcpyAsync(&bad[1], d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
        AHOP_HIP(hipStreamSynchronize(s));
        if (bad[1] == 0) {
            hipLaunchKernelGGL(k_fill_m1, dim3(b1), dim3(256), 0, s, d_slot, n);
            hipLaunchKernelGGL(k_inverse_order, dim3(b1), dim3(256), 0, s, d_order, n, d_slot);
            hipLaunchKernelGGL(k_count_negative, dim3(b1), dim3(256), 0, s, d_slot, n, d_bad);
            AHOP_HIP(hipMemcpyAsync(&bad[1], d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
        }
    }
    AHOP_HIP(hipStreamSynchronize(s));
    if (bad[0]) throw std::runtime_error(std::to_string(bad[0]) + " neighbour ids outside [0, npart)");
    if (bad[1]) throw std::runtime_error("order is not a permutation of [0, npart)");
    if (e0) AHOP_HIP(hipEventRecord(e0, s));
    const unsigned b3 = (unsigned)((3 * n + 255) / 256);
    hipLaunchKernelGGL(k_zero, dim3(b3), dim3(256), 0, s, d_gpos, 3 * n);
    hipLaunchKernelGGL(k_zero, dim3(b1), dim3(256), 0, s, d_gmass, n);
    if (d_order)
        hipLaunchKernelGGL(k_sph_adjoint<true>, dim3((unsigned)((n + ADJ_BS - 1) / ADJ_BS)),
                           dim3(ADJ_BS), 0, s, d_pos, d_mass, d_nbr, d_order, d_slot, d_grad_rho, n,
                           K, Lx, Ly, Lz, A, d_gpos, d_gmass);
    else
        hipLaunchKernelGGL(k_sph_adjoint<false>, dim3((unsigned)((n + ADJ_BS - 1) / ADJ_BS)),
                           dim3(ADJ_BS), 0, s, d_pos, d_mass, d_nbr, nullptr, nullptr, d_grad_rho,
                           n, K, Lx, Ly, Lz, A, d_gpos, d_gmass);
    if (e1) AHOP_HIP(hipEventRecord(e1, s));
    AHOP_HIP(hipStreamSynchronize(s));
}

// Device memory of the allocation-free entry point: the error counter (first 256 bytes) and
// the inverse of the order (ADJOINT_NOALLOC_MAX int32).  Every GPU holds its own copy for the
// life of the process; host code reaches it through the address hipGetSymbolAddress gives on
// that device, never through the symbol itself.
constexpr int64_t ADJOINT_NOALLOC_MAX = 4194304;
constexpr int64_t SCRATCH_HEAD = 64;   // int32 entries before the inverse order (256 bytes)
__device__ __align__(256) int32_t scratch[SCRATCH_HEAD + ADJOINT_NOALLOC_MAX];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
int32_t *scratch_ptr[MAX_DEVICES] = {};

// the current device's address of `scratch`, looked up once per device
int32_t *device_scratch()
{
    int dev = 0;
    AHOP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        AHOP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("no device address for the adjoint's scratch");
        scratch_ptr[dev] = static_cast<int32_t *>(p);
    }
    return scratch_ptr[dev];
}

}  // namespace

float sph_density_adjoint(const double *d_pos, const double *d_mass, const int32_t *d_nbr,
                          const int32_t *d_order, const double *d_grad_rho, int64_t n, int K,
                          double Lx, double Ly, double Lz, double *d_gpos, double *d_gmass,
                          hipStream_t s)
{
    struct Scratch {
        int *bad = nullptr;
        int32_t *slot = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Scratch()
        {
            if (bad) (void)hipFree(bad);
            if (slot) (void)hipFree(slot);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } w;
    check_args(n, K, Lx, Ly, Lz);
    AHOP_HIP(hipMalloc(&w.bad, sizeof(int)));
    if (d_order) AHOP_HIP(hipMalloc(&w.slot, sizeof(int32_t) * n));
    AHOP_HIP(hipEventCreate(&w.e0));
    AHOP_HIP(hipEventCreate(&w.e1));
    adjoint_launch(d_pos, d_mass, d_nbr, d_order, d_grad_rho, n, K, Lx, Ly, Lz, d_gpos, d_gmass,
                   w.bad, w.slot, s, w.e0, w.e1);
    float ms = 0;
    AHOP_HIP(hipEventElapsedTime(&ms, w.e0, w.e1));
    return ms;
}

void sph_density_adjoint_noalloc(const double *d_pos, const double *d_mass,
                                 const int32_t *d_nbr, const int32_t *d_order,
                                 const double *d_grad_rho, int64_t n, int K, double Lx, double Ly,
                                 double Lz, double *d_gpos, double *d_gmass, hipStream_t s)
{
    check_args(n, K, Lx, Ly, Lz);
    if (n > ADJOINT_NOALLOC_MAX)
        throw std::runtime_error("the allocation-free adjoint takes at most " +
                                 std::to_string(ADJOINT_NOALLOC_MAX) + " particles (got " +
                                 std::to_string(n) + ")");
    int32_t *base = device_scratch();
    adjoint_launch(d_pos, d_mass, d_nbr, d_order, d_grad_rho, n, K, Lx, Ly, Lz, d_gpos, d_gmass,
                   reinterpret_cast<int *>(base), base + SCRATCH_HEAD, s, nullptr, nullptr);
}

}  // namespace ahop
