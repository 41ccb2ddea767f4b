// ahop_kernels.hip -- MI355X (gfx950) kernels of the AdaptaHOP hot path.
//
// Data layout in HBM (N particles):
//   key[N] u64 Morton codes -> radix sort -> perm[N] (sorted slot -> original index)
//   spos[N] double4 (x, y, z, mass) in sorted order (32 B, one aligned load per particle)
//   binary radix tree over the sorted codes: first/last/left/right[N-1] int32,
//   box[6*(N-1)] double (lo xyz, hi xyz) per internal node
//   rho_s[N] f64, nbr[N*nhop] i32 (sorted slots of the first nhop neighbours)
//
// Reference semantics reproduced (fortran/compute_adaptahop.f90):
//   kNN (walk_tree2 :2867-2962): self excluded by index, d2 = (dx*dx + dy*dy) + dz*dz with
//     dx = min(|dx|, L-|dx|), k smallest kept (strict <)
//   density (omp_compute_density :2816-2838, spline :3270-3287): sequential sum over the
//     first k-1 neighbours in ascending distance, single-rounded pi
//   hop (find_local_maxima :774-795): densest of {i} U first nhop neighbours, ties -> min index
//   group ids (:797-861): root of the hop forest, roots numbered by ascending particle index
//   saddle list (dev_compute_saddle_list :2433-2813): per (g1,g2) max over member pairs of
//     min(rho_i, rho_j), kept when both directions exist, value = min of the two directions
//
// All FP64 arithmetic that decides a reference-visible value is compiled without
// contraction (-ffp-contract=off) so it rounds exactly like the reference on x86-64.
#include "ahop_common.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

namespace ahop {

static thread_local std::string g_err;
const char *last_error() { return g_err.c_str(); }
void set_error(const std::string &m) { g_err = m; }

void *DevBuf::ensure(size_t nbytes)
{
    if (nbytes <= bytes && p) return p;
    release();
    size_t nb = nbytes < 256 ? 256 : nbytes;
    AHOP_HIP(hipMalloc(&p, nb));
    bytes = nb;
    return p;
}
void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

// single-rounded pi of the reference (f90:58: default-real literal in a real(8) parameter)
__device__ __constant__ double PI_REF = 3.1415927410125732421875;

constexpr int WAVE = 64;
constexpr int STACK = 128;
constexpr float U32 = 5.9604644775390625e-8f;   // 2^-24, unit roundoff of binary32

// ------------------------------------------------------------------------------------------
// small device helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t spread21(uint64_t x)
{
    x &= 0x1fffffULL;
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

__device__ __forceinline__ double wrapd(double d, double L)
{
    double h = 0.5 * L;
    if (d >= h) return d - L;
    if (d < -h) return d + L;
    return d;
}

// spline f90:3270-3287
__device__ __forceinline__ double spline_ref(double x)
{
    if (x <= 1.0) {
        double x2 = x * x;
        double x3 = x2 * x;
        return 1.0 - 1.5 * x2 + 0.75 * x3;
    } else if (x <= 2.0) {
        double x2 = 2.0 - x;
        double x3 = x2 * x2 * x2;
        return 0.25 * x3;
    }
    return 0.0;
}

// reference periodic distance of two absolute positions (walk_tree2 :2931-2941)
__device__ __forceinline__ double d2_ref(double ax, double ay, double az, double bx, double by,
                                         double bz, double Lx, double Ly, double Lz)
{
    double dx = fabs(bx - ax);
    dx = fmin(dx, Lx - dx);
    double d = dx * dx;
    double dy = fabs(by - ay);
    dy = fmin(dy, Ly - dy);
    d = d + dy * dy;
    double dz = fabs(bz - az);
    dz = fmin(dz, Lz - dz);
    d = d + dz * dz;
    return d;
}

__device__ __forceinline__ int logical_block(int nblocks)
{
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
    // logical waves so that neighbouring query waves share their L2.
    int b = blockIdx.x;
    int g8 = nblocks / 8, rem = nblocks % 8;
    int x = b % 8;
    return x * g8 + min(x, rem) + b / 8;
}

// a value that is the same on every lane, as a wave-uniform (scalar) value
__device__ __forceinline__ double uniform_d(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}

// ------------------------------------------------------------------------------------------
// tree build
// ------------------------------------------------------------------------------------------
// Morton keys keep the top MORTON_KEY_BITS of the 63-bit interleaving (16 bits per axis at
// 48): the radix sort then takes 6 passes of 8 bits instead of 8.  Particles in one cell of
// 2^-16 of the box share a key and keep their index order; the tree treats them as
// duplicates, and the results do not depend on the tree's shape.
#ifndef MORTON_KEY_BITS
#define MORTON_KEY_BITS 48
#endif
__global__ void k_morton(const double *__restrict__ pos, int64_t n, double Lx, double Ly, double Lz,
                         uint64_t *__restrict__ key, int32_t *__restrict__ iota, int *__restrict__ bad)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    bool ok = (x >= -0.5 * Lx) && (x < 0.5 * Lx) && (y >= -0.5 * Ly) && (y < 0.5 * Ly) &&
              (z >= -0.5 * Lz) && (z < 0.5 * Lz);
    if (!ok) atomicAdd(bad, 1);
    const double S = 2097152.0;
    double fx = (x / Lx + 0.5) * S, fy = (y / Ly + 0.5) * S, fz = (z / Lz + 0.5) * S;
    uint64_t ix = (uint64_t)fmin(fmax(fx, 0.0), S - 1.0);
    uint64_t iy = (uint64_t)fmin(fmax(fy, 0.0), S - 1.0);
    uint64_t iz = (uint64_t)fmin(fmax(fz, 0.0), S - 1.0);
    key[i] = ((spread21(ix) << 2) | (spread21(iy) << 1) | spread21(iz)) >> (63 - MORTON_KEY_BITS);
    iota[i] = (int32_t)i;
}

__global__ void k_gather(const double *__restrict__ pos, const double *__restrict__ mass,
                         const int32_t *__restrict__ perm, int64_t n, double4 *__restrict__ spos)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    int64_t o = perm[s];
    spos[s] = make_double4(pos[3 * o], pos[3 * o + 1], pos[3 * o + 2], mass[o]);
}

__device__ __forceinline__ int delta(const uint64_t *__restrict__ k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

// Binary radix tree (Karras 2012) over sorted keys; internal node i in [0, n-2].
__global__ void k_radix_tree(const uint64_t *__restrict__ key, int n, int32_t *__restrict__ first,
                             int32_t *__restrict__ last, int32_t *__restrict__ left,
                             int32_t *__restrict__ right, int32_t *__restrict__ parent,
                             int32_t *__restrict__ leaf_parent)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(key, n, i, i + 1) - delta(key, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = delta(key, n, i, i - d);
    int lmax = 2;
    while (delta(key, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(key, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(key, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(key, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int f = min(i, j), la = max(i, j);
    int lc = (f == gamma) ? ~gamma : gamma;
    int rc = (la == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    first[i] = f;
    last[i] = la;
    left[i] = lc;
    right[i] = rc;
    if (lc >= 0) parent[lc] = i; else leaf_parent[~lc] = i;
    if (rc >= 0) parent[rc] = i; else leaf_parent[~rc] = i;
    if (i == 0) parent[0] = -1;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ void child_box(int c, const double4 *__restrict__ spos,
                                          const double *box, double b[6])
{
    if (c < 0) {
        double4 p = spos[~c];
        b[0] = b[3] = p.x;
        b[1] = b[4] = p.y;
        b[2] = b[5] = p.z;
    } else {
        for (int a = 0; a < 6; ++a) b[a] = box[6 * (int64_t)c + a];
    }
}

// FP32 copy of a box for the traversal tests: centre rounded to nearest, half-widths rounded
// up (the test's margin covers the centre rounding); a particle's box (lo = hi) is its
// FP32-rounded position with zero half-widths
__device__ __forceinline__ void fp32_box(const double lo[3], const double hi[3], float4 &a,
                                         float4 &b)
{
    float c[3], h[3];
    for (int q = 0; q < 3; ++q) {
        c[q] = (float)(0.5 * (lo[q] + hi[q]));
        h[q] = __double2float_ru(0.5 * (hi[q] - lo[q]));
    }
    a = make_float4(c[0], c[1], c[2], h[0]);
    b = make_float4(h[1], h[2], 0.0f, 0.0f);
}

__device__ __forceinline__ void store_box(int node, const double lo[3], const double hi[3],
                                          double *box, float4 *fbox)
{
    for (int a = 0; a < 3; ++a) {
        box[6 * (int64_t)node + a] = lo[a];
        box[6 * (int64_t)node + a + 3] = hi[a];
    }
    float4 a, b;
    fp32_box(lo, hi, a, b);
    fbox[2 * (int64_t)node] = a;
    fbox[2 * (int64_t)node + 1] = b;
}

// the visit record's part for child k of a node: the child's FP32 box and its slot range
// [cf, cl] (ints as float bits)
__device__ __forceinline__ void store_child_rec(int node, int k, const double lo[3],
                                                const double hi[3], int cf, int cl,
                                                float4 *cbox)
{
    float4 a, b;
    fp32_box(lo, hi, a, b);
    cbox[4 * (int64_t)node + 2 * k] = a;
    cbox[4 * (int64_t)node + 2 * k + 1] =
        make_float4(b.x, b.y, __int_as_float(cf), __int_as_float(cl));
}

// Bottom-up bounding boxes in three passes, each merging a node once both children are done
// and writing its FP64 box, its FP32 copy and, in the node's visit record, both children's
// FP32 boxes and slot ranges.  A fence per climbing thread at device scope (needed for
// cross-XCD visibility) costs ~1e2 ms at 16.8M particles, so no pass climbs across
// workgroups:
//  pass 1 (k_refit_local): one workgroup per 256 consecutive leaves merges every node whose
//    sorted range lies inside its 256-leaf block.  Such a node's index lies in the block too
//    (an internal node is one end of its own range), so the block preloads the links of its
//    node indices and its leaves' positions into LDS and merges boxes there, communicating
//    through LDS counters and workgroup-scope fences only;
//  pass 2 (k_refit_mid): one workgroup per 65 536 leaves lists the nodes whose range crosses a
//    256-leaf block but lies inside its own block, and merges them in sweeps separated by
//    barriers: a node merges in a sweep when both children were done before it (a leaf, a
//    pass-1 node, or a node merged in an earlier sweep, whose writes the barrier made
//    visible).  Nodes that cross its block go to a list for
//  pass 3 (k_refit_top): one workgroup climbing that list bottom-up with counters (sweeps
//    over it took 1.0 ms at 16.8M particles: the top of the tree is deep).
// Pass 2 stamps a merged node's flag with the sweep's number, so that a child merged in the
// current sweep is never taken as done.
constexpr int REFIT_BS = 256;
constexpr int MID_S = 65536;     // leaves per pass-2 workgroup
constexpr int MID_BS = 1024;
constexpr int MID_CAP = 4096;    // pass-2 list per workgroup (more go to pass 3)
constexpr int TOP_STAMP = 1 << 20;

// the FP64 box and slot range of child c (a particle for c < 0)
__device__ __forceinline__ void child_box_range(int c, const double4 *__restrict__ spos,
                                                const int32_t *__restrict__ first,
                                                const int32_t *__restrict__ last,
                                                const double *box, double lo[3], double hi[3],
                                                int &cf, int &cl)
{
    double b[6];
    child_box(c, spos, box, b);
    for (int q = 0; q < 3; ++q) {
        lo[q] = b[q];
        hi[q] = b[q + 3];
    }
    if (c < 0) {
        cf = cl = ~c;
    } else {
        cf = first[c];
        cl = last[c];
    }
}

__device__ __forceinline__ void merge_node(int node, const double4 *__restrict__ spos,
                                           const int32_t *__restrict__ first,
                                           const int32_t *__restrict__ last,
                                           const int32_t *__restrict__ left,
                                           const int32_t *__restrict__ right, double *box,
                                           float4 *fbox, float4 *cbox, int4 *nrec)
{
    // the links and range of the node's visit record
    nrec[node] = make_int4(left[node], right[node], first[node], last[node]);
    double alo[3], ahi[3], blo[3], bhi[3], lo[3], hi[3];
    int af, al, bf, bl;
    child_box_range(left[node], spos, first, last, box, alo, ahi, af, al);
    child_box_range(right[node], spos, first, last, box, blo, bhi, bf, bl);
    for (int q = 0; q < 3; ++q) {
        lo[q] = fmin(alo[q], blo[q]);
        hi[q] = fmax(ahi[q], bhi[q]);
    }
    store_box(node, lo, hi, box, fbox);
    store_child_rec(node, 0, alo, ahi, af, al, cbox);
    store_child_rec(node, 1, blo, bhi, bf, bl, cbox);
}

// Pass 1 writes of a node only what is read later: its FP64 box when its parent is merged by
// pass 2 or 3 (which read children's boxes from memory), its FP32 box when a traversal can
// push it (the root, or a child of a node of more than `bucket` particles, the nodes a
// traversal expands), and its children's records when it holds more than `bucket`
// particles; the nodes inside a bucket-sized subtree, most of the tree, write nothing (writing
// every record for every node took 1.1 ms at 16.8M particles).
__global__ __launch_bounds__(REFIT_BS) void k_refit_local(
    int n, int bucket, const double4 *__restrict__ spos, const int32_t *__restrict__ first,
    const int32_t *__restrict__ last, const int32_t *__restrict__ left,
    const int32_t *__restrict__ right, const int32_t *__restrict__ parent,
    const int32_t *__restrict__ leaf_parent, double *box, float4 *fbox, float4 *cbox,
    int4 *nrec)
{
    __shared__ int cnt[REFIT_BS];
    __shared__ int lf[REFIT_BS], ll[REFIT_BS], lc0[REFIT_BS], lc1[REFIT_BS], lpar[REFIT_BS];
    __shared__ double lpos[3][REFIT_BS];      // the block's leaves
    __shared__ double lbox[6][REFIT_BS];      // boxes of the block's local nodes
    const int B0 = blockIdx.x * REFIT_BS;
    const int B1 = min(n, B0 + REFIT_BS) - 1;   // last leaf of the block
    const int t = threadIdx.x;
    const int p = B0 + t;
    cnt[t] = 0;
    if (p <= B1) {
        const double4 q = spos[p];
        lpos[0][t] = q.x;
        lpos[1][t] = q.y;
        lpos[2][t] = q.z;
        if (p < n - 1) {
            lf[t] = first[p];
            ll[t] = last[p];
            lc0[t] = left[p];
            lc1[t] = right[p];
            lpar[t] = parent[p];
        }
    }
    __syncthreads();
    if (p > B1) return;
    int node = leaf_parent[p];
    while (node >= B0 && node <= B1 && node < n - 1 && lf[node - B0] >= B0 &&
           ll[node - B0] <= B1) {
        const int u = node - B0;
        __threadfence_block();
        if (atomicAdd(&cnt[u], 1) == 0) return;   // sibling subtree not done yet
        __threadfence_block();
        // children of a local node are the block's leaves or local nodes
        const int c[2] = {lc0[u], lc1[u]};
        double clo[2][3], chi[2][3], lo[3], hi[3];
        for (int k = 0; k < 2; ++k)
            for (int a = 0; a < 3; ++a) {
                clo[k][a] = c[k] < 0 ? lpos[a][~c[k] - B0] : lbox[a][c[k] - B0];
                chi[k][a] = c[k] < 0 ? lpos[a][~c[k] - B0] : lbox[a + 3][c[k] - B0];
            }
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(clo[0][a], clo[1][a]);
            hi[a] = fmax(chi[0][a], chi[1][a]);
            lbox[a][u] = lo[a];
            lbox[a + 3][u] = hi[a];
        }
        const int pn = lpar[u];
        const bool par_local = pn >= B0 && pn <= B1 && pn < n - 1 && lf[pn - B0] >= B0 &&
                               ll[pn - B0] <= B1;
        const bool par_big = !par_local || ll[pn - B0] - lf[pn - B0] + 1 > bucket;
        if (!par_local)
            for (int a = 0; a < 3; ++a) {
                box[6 * (int64_t)node + a] = lo[a];
                box[6 * (int64_t)node + a + 3] = hi[a];
            }
        if (par_big) {
            nrec[node] = make_int4(lc0[u], lc1[u], lf[u], ll[u]);
            float4 fa, fb;
            fp32_box(lo, hi, fa, fb);
            fbox[2 * (int64_t)node] = fa;
            fbox[2 * (int64_t)node + 1] = fb;
        }
        if (ll[u] - lf[u] + 1 > bucket)
            for (int k = 0; k < 2; ++k) {
                const int cf = c[k] < 0 ? ~c[k] : lf[c[k] - B0];
                const int cl = c[k] < 0 ? ~c[k] : ll[c[k] - B0];
                store_child_rec(node, k, clo[k], chi[k], cf, cl, cbox);
            }
        node = pn;
    }
}

// a child is done for the sweep `stamp` when it is a particle, a pass-1 node, or was merged
// with a smaller stamp
__device__ __forceinline__ bool child_done(int c, const int32_t *__restrict__ first,
                                           const int32_t *__restrict__ last,
                                           const int32_t *flag, int stamp)
{
    if (c < 0) return true;
    if (first[c] / REFIT_BS == last[c] / REFIT_BS) return true;
    const int f = flag[c];
    return f != 0 && f < stamp;
}

// Merges the listed nodes (list[k] < 0: already merged) in sweeps, stamps base + 1,
// base + 2, ..., until all are merged or a sweep merges none (a child that this workgroup
// cannot merge: the caller hands the rest on).  Run by every thread of the workgroup; each
// list entry is read and written by one thread only.  Returns whether nodes are left.
__device__ __forceinline__ bool refit_sweeps(int32_t *list, int m, int base, int *more,
                                             int *progress, const double4 *__restrict__ spos,
                                             const int32_t *__restrict__ first,
                                             const int32_t *__restrict__ last,
                                             const int32_t *__restrict__ left,
                                             const int32_t *__restrict__ right, int32_t *flag,
                                             double *box, float4 *fbox, float4 *cbox, int4 *nrec)
{
    for (int sweep = 1;; ++sweep) {
        const int stamp = base + sweep;
        if (threadIdx.x == 0) {
            *more = 0;
            *progress = 0;
        }
        __syncthreads();
        bool pending = false, merged = false;
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            const int i = list[k];
            if (i < 0) continue;
            if (child_done(left[i], first, last, flag, stamp) &&
                child_done(right[i], first, last, flag, stamp)) {
                merge_node(i, spos, first, last, left, right, box, fbox, cbox, nrec);
                flag[i] = stamp;
                list[k] = -1;
                merged = true;
            } else {
                pending = true;
            }
        }
        if (pending) *more = 1;
        if (merged) *progress = 1;
        __syncthreads();
        const int again = *more, moved = *progress;
        __syncthreads();
        if (!again) return false;
        if (!moved) return true;
    }
}

__global__ __launch_bounds__(MID_BS) void k_refit_mid(
    int n, const double4 *__restrict__ spos, const int32_t *__restrict__ first,
    const int32_t *__restrict__ last, const int32_t *__restrict__ left,
    const int32_t *__restrict__ right, int32_t *flag, double *box, float4 *fbox, float4 *cbox,
    int4 *nrec, int32_t *top, int *ntop)
{
    __shared__ int32_t list[MID_CAP];
    __shared__ int nlist, more, progress;
    const int B0 = blockIdx.x * MID_S;
    const int B1 = min(n, B0 + MID_S) - 1;
    if (threadIdx.x == 0) nlist = 0;
    __syncthreads();
    for (int i = B0 + threadIdx.x; i <= min(B1, n - 2); i += MID_BS) {
        const int f = first[i], l = last[i];
        if (f / REFIT_BS == l / REFIT_BS) continue;   // merged by pass 1
        if (f >= B0 && l <= B1) {
            const int k = atomicAdd(&nlist, 1);
            if (k < MID_CAP) {
                list[k] = i;
                continue;
            }
        }
        top[atomicAdd(ntop, 1)] = i;                  // crosses this block (or the list is full)
    }
    __syncthreads();
    const int m = min(nlist, MID_CAP);
    if (refit_sweeps(list, m, 0, &more, &progress, spos, first, last, left, right, flag, box,
                     fbox, cbox, nrec)) {
        // nodes above one that went to pass 3 (a full list): pass 3 takes them as well
        for (int k = threadIdx.x; k < m; k += MID_BS)
            if (list[k] >= 0) top[atomicAdd(ntop, 1)] = list[k];
    }
}

// Pass 3: one workgroup.  Each listed node's flag is first set to -(1 + its children not yet
// merged) (a child is merged when it is a particle, a pass-1 node or has a positive pass-2
// stamp), or to a start mark when none is pending; then threads climb from the starts,
// merging, and a parent is
// merged by the thread whose child arrives last (workgroup-scope fences: the whole pass is on
// one CU).  The parent of a listed node is listed too (its range is wider), or is none.
__global__ __launch_bounds__(MID_BS) void k_refit_top(
    const double4 *__restrict__ spos, const int32_t *__restrict__ first,
    const int32_t *__restrict__ last, const int32_t *__restrict__ left,
    const int32_t *__restrict__ right, const int32_t *__restrict__ parent, int32_t *flag,
    double *box, float4 *fbox, float4 *cbox, int4 *nrec, const int32_t *top, const int *ntop,
    int *stalled)
{
    const int m = *ntop;
    auto merged = [&](int c) {
        return c < 0 || first[c] / REFIT_BS == last[c] / REFIT_BS ||
               __hip_atomic_load(&flag[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > 0;
    };
    // a start (no child pending) is marked TOP_START, which arrivals never produce: a node with
    // children pending counts -1 - pending up to -1 as they arrive
    constexpr int TOP_START = -1000;
    for (int k = threadIdx.x; k < m; k += MID_BS) {
        const int i = top[k];
        const int pending = (merged(left[i]) ? 0 : 1) + (merged(right[i]) ? 0 : 1);
        __hip_atomic_store(&flag[i], pending == 0 ? TOP_START : -1 - pending, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    int climbed = 0;
    for (int k = threadIdx.x; k < m; k += MID_BS) {
        int node = top[k];
        if (__hip_atomic_load(&flag[node], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) !=
            TOP_START)
            continue;   // not a start: a child is still pending
        while (true) {
            merge_node(node, spos, first, last, left, right, box, fbox, cbox, nrec);
            ++climbed;
            __hip_atomic_store(&flag[node], TOP_STAMP, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
            node = parent[node];
            if (node < 0) break;
            __threadfence_block();
            const int old = __hip_atomic_fetch_add(&flag[node], 1, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old != -2) break;   // the parent's other child is still pending
            __threadfence_block();
        }
    }
    // every listed node is merged exactly once; fewer merges mean a node never became ready
    __shared__ int total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    atomicAdd(&total, climbed);
    __syncthreads();
    if (threadIdx.x == 0 && total != m) *stalled = 1;
}

// ------------------------------------------------------------------------------------------
// kNN + SPH density
// ------------------------------------------------------------------------------------------
struct KnnArgs {
    const double4 *spos;
    const int32_t *first, *last, *left, *right;
    const double *box;
    const float4 *fbox;     // per internal node: (cx, cy, cz, hx), (hy, hz, 0, 0) in FP32
    const int4 *nrec;       // per internal node: left, right, first, last
    const float4 *cbox;     // per internal node: both children's boxes and slot ranges
    int n, K, nhop, bucket;
    double Lx, Ly, Lz, xyzl, rho_thr;
    const int32_t *qlist;   // optional explicit queries (sorted slots); null -> all slots
    int nq;
    int qper;               // queries per wave: WAVE, or 1 for a short list of scattered queries
    double *rho_s;
    int32_t *nbr;
    int32_t *fb_list;       // mixed path: queries that need the FP64 path
    int32_t *fb_count;
    unsigned long long *nact;
    int *overflow;
    int nblocks;
    unsigned long long *stats;   // optional: per-launch work counters (see KNN_STAT_*)
};

// work counters accumulated by k_knn when KnnArgs::stats is set
enum { KNN_STAT_NODES = 0, KNN_STAT_BUCKETS, KNN_STAT_CANDIDATES, KNN_STAT_INSERT_STEPS,
       KNN_STAT_INSERTS, KNN_STAT_WAVES, KNN_STAT_LOG_MAX, KNN_STAT_LOG_OVERFLOW,
       KNN_STAT_MISORDER, KNN_STAT_N };

// Heap of the KH smallest keys per lane, lane-interleaved in LDS (entry e of lane l at e*64+l,
// so any per-lane entry index is bank-conflict free).
// 8-ary max-heap: at most 2 levels below the root for up to 73 entries, and each level
// issues all of its children's key and id loads at once, so a replacement costs two
// dependent LDS round trips instead of the ~12 of a binary heap with the same capacity.
constexpr int HEAP_D = 8;

// Place (d, id) at heap node e (the root by default) and sift it down a max-heap of `size`
// entries (HEAP_D-ary, lane-interleaved in LDS).  CHECKED = false reads all HEAP_D children
// of an internal node without bounds checks: the caller keeps the entries past `size` up to
// the last child of the last internal node allocated and at -inf, so they never win.
template <typename KeyT, bool CHECKED = true>
__device__ __forceinline__ void heap_sift(KeyT *hk, int32_t *hi, int lane, int size, KeyT d,
                                          int32_t id, int e = 0)
{
    while (true) {
        const int c0 = HEAP_D * e + 1;
        if (c0 >= size) break;
        KeyT kk[HEAP_D];
        int32_t ii[HEAP_D];
#pragma unroll
        for (int t = 0; t < HEAP_D; ++t) {
            const int c = (!CHECKED || c0 + t < size) ? c0 + t : c0;
            kk[t] = hk[c * WAVE + lane];
            ii[t] = hi[c * WAVE + lane];
        }
        KeyT kc = kk[0];
        int32_t idc = ii[0];
        int ic = c0;
#pragma unroll
        for (int t = 1; t < HEAP_D; ++t)
            if ((!CHECKED || c0 + t < size) && kk[t] > kc) {
                kc = kk[t];
                idc = ii[t];
                ic = c0 + t;
            }
        if (!(kc > d)) break;
        hk[e * WAVE + lane] = kc;
        hi[e * WAVE + lane] = idc;
        e = ic;
    }
    hk[e * WAVE + lane] = d;
    hi[e * WAVE + lane] = id;
}

// Heap of the mixed kernel: packed (FP32 key bits, id) entries, the top two levels in
// registers.  The root (the largest key kept) is entry 0 (k0, i0); its RC children are
// entries 1..RC (rk, ri); every register child c has LC = 8 children, entries
// LC * c + RT - LC + t (t < LC, RT = RC + 1), which are leaves and live in LDS at
// hp[(e - RT) * WAVE + lane].  Up to RT + LC * RC entries fit in these three levels.
// A replacement at the root finds the largest register child without LDS and, only when
// that child has children, reads them from LDS in one round trip.  Up to k = 64 the root
// has RC = 8 children: the k + 1 = 65 entries take 56 LDS entries per lane (5 waves per CU
// by LDS) and 7 of the 8 register children have children.  With 16 register children
// (48 LDS entries, 6 waves per CU) every replacement took a maximum over 16 entries and
// their write-back, and the kernel needed 255 registers and spilled: at k = 64 it issued 145
// thousand VALU instructions per wave against 112 thousand with 8, and ran 99.3 ms against
// 98.6 ms at the 16.8M-particle box (the spare registers then carry two pending candidates
// per lane through the seeding's flushes too: 96.8 ms).  Above k = 64 the root keeps 16
// children (129 entries need them).  The register vectors are indexed with compile-time
// indices only (selects).
template <int KCAP>
constexpr int heap_rc() { return KCAP > 64 ? 16 : 8; }
constexpr int LC = 8;

// first child of register entry c (1 <= c <= RC)
template <int RC>
__device__ __forceinline__ int rt_child0(int c) { return LC * c + RC + 1 - LC; }

// vector types rather than arrays: element reads and writes are register inserts and
// extracts, never loads from a stack slot, so the compiler cannot turn a select over
// entries into an indexed memory access
template <int RC>
struct RegTop {
    typedef float rtk_t __attribute__((ext_vector_type(RC)));
    typedef uint32_t rti_t __attribute__((ext_vector_type(RC)));
    float k0;
    uint32_t i0;
    rtk_t rk;
    rti_t ri;
};

template <int RC>
__device__ __forceinline__ float rt_key(const RegTop<RC> &h, const uint2 *hp, int lane, int e)
{
    constexpr int RT = RC + 1;
    if (e >= RT) return __uint_as_float(hp[(e - RT) * WAVE + lane].x);
    float v = h.k0;
#pragma unroll
    for (int t = 0; t < RC; ++t) v = (e == t + 1) ? h.rk[t] : v;
    return v;
}

template <int RC>
__device__ __forceinline__ uint32_t rt_id(const RegTop<RC> &h, const uint2 *hp, int lane, int e)
{
    constexpr int RT = RC + 1;
    if (e >= RT) return hp[(e - RT) * WAVE + lane].y;
    uint32_t v = h.i0;
#pragma unroll
    for (int t = 0; t < RC; ++t) v = (e == t + 1) ? h.ri[t] : v;
    return v;
}

template <int RC>
__device__ __forceinline__ void rt_set(RegTop<RC> &h, uint2 *hp, int lane, int e, float k,
                                       uint32_t id)
{
    constexpr int RT = RC + 1;
    if (e >= RT) {
        hp[(e - RT) * WAVE + lane] = make_uint2(__float_as_uint(k), id);
    } else {
        h.k0 = e == 0 ? k : h.k0;
        h.i0 = e == 0 ? id : h.i0;
#pragma unroll
        for (int t = 0; t < RC; ++t) {
            const bool m = e == t + 1;
            h.rk[t] = m ? k : h.rk[t];
            h.ri[t] = m ? id : h.ri[t];
        }
    }
}

// sift (d, id) down from register entry c, 1 <= c <= RC: the largest of its LDS children
// (read all LC of them unchecked: the caller keeps entries past `size` at -inf) moves up
// if it exceeds d, and d takes its place
template <bool CHECKED, int RC>
__device__ __forceinline__ void rt_sift_l1(RegTop<RC> &h, uint2 *hp, int lane, int size, float d,
                                           uint32_t id, int c)
{
    constexpr int RT = RC + 1;
    const int c0 = rt_child0<RC>(c);
    float vk = d;
    uint32_t vi = id;
    if (c0 < size) {
        // one base address and constant offsets for the unchecked reads
        const uint2 *base = hp + (c0 - RT) * WAVE + lane;
        uint2 ch[LC];
#pragma unroll
        for (int t = 0; t < LC; ++t) ch[t] = base[((!CHECKED || c0 + t < size) ? t : 0) * WAVE];
        float kc;
        uint32_t idc;
        int tc;
        // the maximum by max instructions, then the first entry equal to it (the same entry
        // as a strict-greater scan): no chain of dependent compare-select pairs.  Checked,
        // the entries past the heap size count as -inf (the first child is inside it).
        float kt[LC];
#pragma unroll
        for (int t = 0; t < LC; ++t)
            kt[t] = (!CHECKED || c0 + t < size) ? __uint_as_float(ch[t].x) : -INFINITY;
        kc = kt[0];
#pragma unroll
        for (int t = 1; t < LC; ++t) kc = fmaxf(kc, kt[t]);
        tc = LC - 1;
        idc = ch[LC - 1].y;
#pragma unroll
        for (int t = LC - 2; t >= 0; --t) {
            const bool m = kt[t] == kc;
            tc = m ? t : tc;
            idc = m ? ch[t].y : idc;
        }
        // the id is used on both sides of the test, so key and id come in one 64-bit read
        const bool up = kc > d;
        vk = up ? kc : d;
        vi = up ? idc : id;
        if (up) hp[(c0 + tc - RT) * WAVE + lane] = make_uint2(__float_as_uint(d), id);
    }
#pragma unroll
    for (int t = 0; t < RC; ++t) {
        const bool m = c == t + 1;
        h.rk[t] = m ? vk : h.rk[t];
        h.ri[t] = m ? vi : h.ri[t];
    }
}

// sift (d, id) down from the root
template <bool CHECKED, int RC>
__device__ __forceinline__ void rt_sift_root(RegTop<RC> &h, uint2 *hp, int lane, int size, float d,
                                             uint32_t id)
{
    float kc;
    uint32_t idc;
    int ic;
    // the maximum by max instructions, then the first entry equal to it; checked, the
    // entries past the heap size count as -inf (with none left, kc = -inf and d stays)
    float kt[RC];
#pragma unroll
    for (int t = 0; t < RC; ++t) kt[t] = (!CHECKED || t + 1 < size) ? h.rk[t] : -INFINITY;
    kc = kt[0];
#pragma unroll
    for (int t = 1; t < RC; ++t) kc = fmaxf(kc, kt[t]);
    ic = RC;
    idc = h.ri[RC - 1];
#pragma unroll
    for (int t = RC - 2; t >= 0; --t) {
        const bool m = kt[t] == kc;
        ic = m ? t + 1 : ic;
        idc = m ? h.ri[t] : idc;
    }
    if (kc > d) {
        h.k0 = kc;
        h.i0 = idc;
        rt_sift_l1<CHECKED, RC>(h, hp, lane, size, d, id, ic);
    } else {
        h.k0 = d;
        h.i0 = id;
    }
}

// Heap entry access for both kernels: register-top packed (mixed) or key and id arrays
// (FP64).
template <bool PACKED, typename KeyT, int RC>
__device__ __forceinline__ KeyT heap_key(const RegTop<RC> &h, const uint2 *hp, const KeyT *hk,
                                         int lane, int e)
{
    if constexpr (PACKED) return rt_key(h, hp, lane, e);
    else return hk[e * WAVE + lane];
}
template <bool PACKED, typename KeyT, int RC>
__device__ __forceinline__ int32_t heap_id(const RegTop<RC> &h, const uint2 *hp, const int32_t *hi,
                                           int lane, int e)
{
    if constexpr (PACKED) return (int32_t)rt_id(h, hp, lane, e);
    else return hi[e * WAVE + lane];
}
template <bool PACKED, typename KeyT, int RC>
__device__ __forceinline__ void heap_set(RegTop<RC> &h, uint2 *hp, KeyT *hk, int32_t *hi, int lane,
                                         int e, KeyT k, int32_t id)
{
    if constexpr (PACKED) {
        rt_set(h, hp, lane, e, (float)k, (uint32_t)id);
    } else {
        hk[e * WAVE + lane] = k;
        hi[e * WAVE + lane] = id;
    }
}
// sift from entry e: the root, a register child, or (FP64) any entry; the mixed kernel's
// LDS entries are leaves and are never sifted from
template <bool PACKED, bool CHECKED, typename KeyT, int RC>
__device__ __forceinline__ void heap_sift_any(RegTop<RC> &h, uint2 *hp, KeyT *hk, int32_t *hi,
                                              int lane, int size, KeyT d, int32_t id, int e)
{
    if constexpr (PACKED) {
        if (e == 0) rt_sift_root<CHECKED, RC>(h, hp, lane, size, (float)d, (uint32_t)id);
        else rt_sift_l1<CHECKED, RC>(h, hp, lane, size, (float)d, (uint32_t)id, e);
    } else {
        heap_sift<KeyT, CHECKED>(hk, hi, lane, size, d, id, e);
    }
}

#ifndef AHOP_FP64_NBUF
#define AHOP_FP64_NBUF 2
#endif
#ifndef AHOP_FP64_FILL
#define AHOP_FP64_FILL 1
#endif
#ifndef AHOP_WAVE_SEED
#define AHOP_WAVE_SEED 1
#endif
#ifndef AHOP_SEED_PAD
#define AHOP_SEED_PAD 16
#endif
// Staged buckets of the mixed kernel.  LDS is allocated per wave in granules of 1280 B
// (160 KB per CU): 21 granules (26880 B) are the budget for 6 waves per CU, 25 for 5.  At
// k = 64 the LDS heap (48 x 64 x 8 B), the stack (512 B) and 2 buckets of x, y, z (1536 B)
// take 26624 B, 6 waves; a 3rd bucket fits only 5 and measured 25% slower (at C3,
// 16.8 M particles: 138 ms against 172 ms).
// Pending candidates each lane takes per step of the mixed kernel's traversal flushes: two,
// whose stage reads are in flight together, inserted smallest first (C3 kernel 106.8 ->
// 104.5 ms; four took 108.4 ms, and two in the seeding flushes too 106.0 ms: registers).
#ifndef AHOP_PICKW
#define AHOP_PICKW 2
#endif
#ifndef AHOP_MIXED_NBUF
#define AHOP_MIXED_NBUF 2
#endif
#include "knn_wave.h"
#include "knn_sel.h"

// a double held by lane i, as a wave-uniform value
__device__ __forceinline__ double lane_d(double v, int i)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, i);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), i);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// FP64 search of one listed query per wave (the mixed path's uncertified queries), the
// wave's lanes sharing it: lane i holds the i-th nearest candidate kept so far, ascending,
// a bucket's candidates are evaluated one per lane, and each accepted candidate (FP64 key
// below the k-th, f90:2934/:2943) is inserted at its rank, after the kept entries with an
// equal key (:2945).  Nodes are visited nearest first and tested again, against the
// tightened bound, when popped.  Then the density in the reference's order and arithmetic
// (omp_compute_density f90:2816-2838).  K <= 64.
__global__ __launch_bounds__(64) void k_knn_one(KnnArgs A)
{
    __shared__ int32_t snode[STACK];
    __shared__ float sbd[STACK];
    const int lane = threadIdx.x;
    const int w = blockIdx.x;
    if (w >= A.nq) return;
    const int q = A.qlist[w];
    const int K = A.K;
    const double4 P = A.spos[q];
    const double Lx = A.Lx, Ly = A.Ly, Lz = A.Lz;
    const float pxf = (float)P.x, pyf = (float)P.y, pzf = (float)P.z;
    const float Lxf = (float)Lx, Lyf = (float)Ly, Lzf = (float)Lz;
    const float mLx = 4.0f * U32 * Lxf, mLy = 4.0f * U32 * Lyf, mLz = 4.0f * U32 * Lzf;
    double key = INFINITY;    // lanes >= K hold nothing
    int32_t id = -1;
    double tau = INFINITY;    // the K-th key
    const bool kept = lane < K;
    int sp = 0;
    bool overflow = false;
    if (lane == 0) {
        snode[0] = 0;
        sbd[0] = 0.0f;
    }
    sp = 1;
    __syncthreads();
    while (sp > 0) {
        --sp;
        const int node = __builtin_amdgcn_readfirstlane(snode[sp]);
        const float bd = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(sbd[sp])));
        if (!((double)bd < tau)) continue;
        int f, l;
        int4 rec = make_int4(0, 0, 0, 0);
        if (node < 0) {
            f = l = ~node;
        } else {
            rec = A.nrec[node];
            f = rec.z;
            l = rec.w;
        }
        if (node < 0 || l - f + 1 <= A.bucket || (A.n <= A.bucket && node == 0)) {
            const int cnt = l - f + 1;
            double d = INFINITY;
            if (lane < cnt && f + lane != q) {
                const double4 c = A.spos[f + lane];
                d = d2_ref(P.x, P.y, P.z, c.x, c.y, c.z, Lx, Ly, Lz);
            }
            uint64_t m = __ballot(d < tau);
            while (m) {
                const int j = __ffsll((long long)m) - 1;
                m &= m - 1;
                const double c = lane_d(d, j);
                if (!(c < tau)) continue;
                const int rank = __popcll(__ballot(kept && key <= c));
                const double kprev = __shfl_up(key, 1);
                const int32_t iprev = __shfl_up(id, 1);
                if (kept && lane > rank) {
                    key = kprev;
                    id = iprev;
                }
                if (lane == rank) {
                    key = c;
                    id = f + j;
                }
                tau = lane_d(key, K - 1);
            }
            continue;
        }
        const float4 cb0 = A.cbox[4 * node], cb1 = A.cbox[4 * node + 1];
        const float4 cb2 = A.cbox[4 * node + 2], cb3 = A.cbox[4 * node + 3];
        const float bd1 = box_dist2_at(cb0.x, cb0.y, cb0.z, cb0.w, cb1.x, cb1.y, pxf, pyf, pzf,
                                       Lxf, Lyf, Lzf, mLx, mLy, mLz);
        const float bd2 = box_dist2_at(cb2.x, cb2.y, cb2.z, cb2.w, cb3.x, cb3.y, pxf, pyf, pzf,
                                       Lxf, Lyf, Lzf, mLx, mLy, mLz);
        const bool n1 = (double)bd1 < tau, n2 = (double)bd2 < tau;
        if (sp + 2 > STACK) {
            overflow = true;
            break;
        }
        // the nearer child is popped first
        const bool first1 = bd1 <= bd2;
        const int cf = first1 ? rec.x : rec.y, cs = first1 ? rec.y : rec.x;
        const float bf = first1 ? bd1 : bd2, bs = first1 ? bd2 : bd1;
        const bool nf = first1 ? n1 : n2, ns = first1 ? n2 : n1;
        if (ns) {
            if (lane == 0) {
                snode[sp] = cs;
                sbd[sp] = bs;
            }
            ++sp;
        }
        if (nf) {
            if (lane == 0) {
                snode[sp] = cf;
                sbd[sp] = bf;
            }
            ++sp;
        }
        __syncthreads();
    }
    if (overflow) {
        if (lane == 0) atomicAdd(A.overflow, 1);
        return;
    }
    // density from the K kept entries, summed in ascending order
    const double d2k = lane_d(key, K - 1);
    const double r = sqrt(d2k) * 0.5;
    const double r3 = r * r * r;
    const double unsr = 1.0 / r;
    double term = 0.0;
    if (lane < K - 1) term = A.spos[id].w * spline_ref(sqrt(key) * unsr);
    double contrib = 0.0;
    for (int i = 0; i < K - 1; ++i) contrib = contrib + lane_d(term, i);
    if (lane == 0) A.rho_s[q] = A.xyzl * (contrib + P.w) / (PI_REF * r3);
    if (lane < A.nhop && lane < K) A.nbr[(int64_t)q * A.nhop + lane] = id;
}

// ------------------------------------------------------------------------------------------
// hop, group ids
// ------------------------------------------------------------------------------------------
// active-particle count (density > rho_threshold, f90:712-715)
// (grid-stride: each wave sums its share and adds it with one atomic, so a few thousand
// atomics serve the whole array; one atomic per wave of a full-size grid, 262 144 on one
// address at 16.8M particles, took 0.9 ms)
__global__ void k_count_active(const double *__restrict__ rho, int64_t n, double thr,
                               unsigned long long *cnt)
{
    unsigned c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        c += rho[i] > thr ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, (unsigned long long)c);
}

__global__ void k_hop(const double *__restrict__ rho, const int32_t *__restrict__ nbr,
                      const int32_t *__restrict__ perm, int64_t n, int nhop, double thr,
                      int32_t *__restrict__ hop)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    double densest = rho[q];
    if (!(densest > thr)) {
        hop[q] = -1;
        return;
    }
    int32_t sel = (int32_t)q;
    int32_t osel = perm[q];
    for (int t = 0; t < nhop; ++t) {
        int32_t j = nbr[q * nhop + t];
        double rj = rho[j];
        if (rj > densest) {
            sel = j;
            densest = rj;
            osel = perm[j];
        } else if (rj == densest) {
            int32_t oj = perm[j];
            if (oj < osel) {
                sel = j;
                osel = oj;
            }
        }
    }
    hop[q] = sel;
}

// one pointer-doubling round: out = in[in]; flags a change
__global__ void k_jump(const int32_t *__restrict__ in, int32_t *__restrict__ out, int64_t n,
                       int *changed)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    int32_t p = in[q];
    int32_t r = p;
    if (p >= 0) {
        r = in[p];
        if (r != p) *changed = 1;
    }
    out[q] = r;
}

__global__ void k_root_flags(const int32_t *__restrict__ hop, const int32_t *__restrict__ perm,
                             int64_t n, int32_t *__restrict__ flag_o)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    flag_o[perm[q]] = (hop[q] == q) ? 1 : 0;
}

__global__ void k_groups(const int32_t *__restrict__ root, const int32_t *__restrict__ hop,
                         const int32_t *__restrict__ perm, const int32_t *__restrict__ rank_o,
                         const double *__restrict__ rho_s, int64_t n, int32_t *__restrict__ gid_s,
                         int32_t *__restrict__ gid_o, int32_t *__restrict__ hop_o,
                         double *__restrict__ rho_o, double *__restrict__ densityg)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    int32_t o = perm[q];
    int32_t r = root[q];
    int32_t g = 0;
    if (r >= 0) {
        g = rank_o[perm[r]] + 1;
        if (r == q) densityg[g - 1] = rho_s[q];
    }
    gid_s[q] = g;
    gid_o[o] = g;
    hop_o[o] = (hop[q] >= 0) ? perm[hop[q]] : -1;
    rho_o[o] = rho_s[q];
}

// ------------------------------------------------------------------------------------------
// saddle edges
// ------------------------------------------------------------------------------------------
__global__ void k_emit_count(const int32_t *__restrict__ gid_s, const int32_t *__restrict__ nbr,
                             int64_t n, int nhop, int32_t *__restrict__ cnt)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    int32_t g1 = gid_s[q];
    int c = 0;
    if (g1 > 0)
        for (int t = 0; t < nhop; ++t) {
            int32_t g2 = gid_s[nbr[q * nhop + t]];
            c += (g2 > 0 && g2 != g1);
        }
    cnt[q] = c;
}

__global__ void k_emit_fill(const int32_t *__restrict__ gid_s, const int32_t *__restrict__ nbr,
                            const double *__restrict__ rho_s, const int32_t *__restrict__ off,
                            int64_t n, int nhop, uint64_t *__restrict__ ekey,
                            double *__restrict__ eval)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    int32_t g1 = gid_s[q];
    if (g1 <= 0) return;
    int32_t o = off[q];
    double r1 = rho_s[q];
    for (int t = 0; t < nhop; ++t) {
        int32_t j = nbr[q * nhop + t];
        int32_t g2 = gid_s[j];
        if (g2 > 0 && g2 != g1) {
            double r2 = rho_s[j];
            ekey[o] = ((uint64_t)(uint32_t)g1 << 32) | (uint32_t)g2;
            eval[o] = r1 < r2 ? r1 : r2;
            ++o;
        }
    }
}

__global__ void k_symmetrize(const uint64_t *__restrict__ ukey, const double *__restrict__ uval,
                             const int32_t *__restrict__ nuniq_p, int32_t *__restrict__ keep,
                             double *__restrict__ sval)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    int nu = *nuniq_p;
    if (u >= nu) return;
    uint64_t k = ukey[u];
    uint64_t rk = (k << 32) | (k >> 32);
    int lo = 0, hi = nu;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (ukey[mid] < rk) lo = mid + 1; else hi = mid;
    }
    bool found = lo < nu && ukey[lo] == rk;
    keep[u] = found ? 1 : 0;
    double a = uval[u];
    sval[u] = found ? (uval[lo] < a ? uval[lo] : a) : a;
}

__global__ void k_compact(const uint64_t *__restrict__ ukey, const double *__restrict__ sval,
                          const int32_t *__restrict__ keep, const int32_t *__restrict__ koff,
                          const int32_t *__restrict__ nuniq_p, uint32_t *__restrict__ g1_out,
                          int32_t *__restrict__ g2_out, double *__restrict__ s_out)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= *nuniq_p || !keep[u]) return;
    int o = koff[u];
    g1_out[o] = (uint32_t)(ukey[u] >> 32);
    g2_out[o] = (int32_t)(uint32_t)(ukey[u] & 0xffffffffu);
    s_out[o] = sval[u];
}

__global__ void k_seg_offsets(const uint32_t *__restrict__ g1, int64_t ne, int ngroups,
                              int32_t *__restrict__ off)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;   // offset for group g+1 (1-based), g <= ngroups
    if (g > ngroups) return;
    uint32_t target = (uint32_t)(g + 1);
    int64_t lo = 0, hi = ne;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (g1[mid] < target) lo = mid + 1; else hi = mid;
    }
    off[g] = (int32_t)lo;
}

// ------------------------------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------------------------------
struct MaxOp {
    __host__ __device__ double operator()(double a, double b) const { return a > b ? a : b; }
};

static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

Engine::Engine()
{
    AHOP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto &e : ev_) AHOP_HIP(hipEventCreate(&e));
}

Engine::~Engine()
{
    release_all();
    for (auto &e : ev_) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::release_all()
{
    DevBuf *all[] = {&b_key, &b_key_tmp, &b_perm, &b_iota, &b_spos, &b_first, &b_last, &b_left,
                     &b_right, &b_parent, &b_leaf_parent, &b_flag, &b_box, &b_temp, &b_rho_s,
                     &b_nbr, &b_hop_s, &b_gid_s, &b_flag_o, &b_rank_o, &b_density, &b_gid, &b_hop,
                     &b_counter, &b_fallback, &b_emit_cnt, &b_emit_off, &b_ekey, &b_eval,
                     &b_ekey2, &b_eval2, &b_ukey, &b_uval, &b_keep, &b_keep_off, &b_skey,
                     &b_sval, &b_sad_off, &b_sad_isad, &b_sad_rho, &b_sad_cnt, &b_densityg,
                     &b_misc, &b_roots, &b_fbox, &b_fb_sorted, &b_nrec, &b_cbox, &b_klog, &b_wctr};
    for (auto *b : all) b->release();
    res_ = DeviceResults{};
}

template <typename T>
static T *as(DevBuf &b, size_t count)
{
    return (T *)b.ensure(sizeof(T) * (count ? count : 1));
}

void Engine::build_tree(const double *d_pos, const double *d_mass, int64_t n, const Params &p)
{
    hipStream_t s = stream_;
    t_.n = n;
    t_.key = as<uint64_t>(b_key, n);
    t_.key_tmp = as<uint64_t>(b_key_tmp, n);
    t_.iota = as<int32_t>(b_iota, n);
    t_.perm = as<int32_t>(b_perm, n);
    t_.spos = as<double4>(b_spos, n);
    int *misc = as<int>(b_misc, 16);
    AHOP_HIP(hipMemsetAsync(misc, 0, 16 * sizeof(int), s));
    hipLaunchKernelGGL(k_morton, dim3(nblk(n)), dim3(256), 0, s, d_pos, n, p.xlong, p.ylong,
                       p.zlong, t_.key_tmp, t_.iota, misc);
    int bad = 0;
    AHOP_HIP(hipMemcpyAsync(&bad, misc, sizeof(int), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    if (bad != 0)
        throw std::runtime_error("positions must be finite and inside [-L/2, L/2) on every axis (" +
                                 std::to_string(bad) + " particles are not)");
    size_t tb = 0;
    AHOP_HIP(rocprim::radix_sort_pairs(nullptr, tb, t_.key_tmp, t_.key, t_.iota, t_.perm,
                                       (size_t)n, 0, MORTON_KEY_BITS, s));
    void *temp = b_temp.ensure(tb);
    AHOP_HIP(rocprim::radix_sort_pairs(temp, tb, t_.key_tmp, t_.key, t_.iota, t_.perm, (size_t)n,
                                       0, MORTON_KEY_BITS, s));
    hipLaunchKernelGGL(k_gather, dim3(nblk(n)), dim3(256), 0, s, d_pos, d_mass, t_.perm, n,
                       t_.spos);
    AHOP_HIP(hipEventRecord(ev_[1], s));
    int64_t ni = n - 1;
    t_.first = as<int32_t>(b_first, ni);
    t_.last = as<int32_t>(b_last, ni);
    t_.left = as<int32_t>(b_left, ni);
    t_.right = as<int32_t>(b_right, ni);
    t_.parent = as<int32_t>(b_parent, ni);
    t_.leaf_parent = as<int32_t>(b_leaf_parent, n);
    t_.flag = as<int32_t>(b_flag, ni);
    t_.box = as<double>(b_box, 6 * ni);
    t_.nrec = as<int4>(b_nrec, ni);
    t_.cbox = as<float4>(b_cbox, 4 * ni);
    hipLaunchKernelGGL(k_radix_tree, dim3(nblk(ni)), dim3(256), 0, s, t_.key, (int)n, t_.first,
                       t_.last, t_.left, t_.right, t_.parent, t_.leaf_parent);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ni)), dim3(256), 0, s, t_.flag, ni, 0);
    t_.fbox = as<float4>(b_fbox, 2 * ni);
    hipLaunchKernelGGL(k_refit_local, dim3(nblk(n, REFIT_BS)), dim3(REFIT_BS), 0, s, (int)n,
                       p.bucket, t_.spos, t_.first, t_.last, t_.left, t_.right, t_.parent, t_.leaf_parent,
                       t_.box, t_.fbox, t_.cbox, t_.nrec);
    // pass 3's list (misc[5] counts it): the roots buffer is free here
    int32_t *top = as<int32_t>(b_roots, n);
    hipLaunchKernelGGL(k_refit_mid, dim3((unsigned)((n + MID_S - 1) / MID_S)), dim3(MID_BS), 0, s,
                       (int)n, t_.spos, t_.first, t_.last, t_.left, t_.right, t_.flag, t_.box,
                       t_.fbox, t_.cbox, t_.nrec, top, misc + 5);
    hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(MID_BS), 0, s, t_.spos, t_.first, t_.last,
                       t_.left, t_.right, t_.parent, t_.flag, t_.box, t_.fbox, t_.cbox, t_.nrec,
                       top, misc + 5, misc + 6);
    AHOP_HIP(hipEventRecord(ev_[2], s));
}

template <int KCAP>
static void launch_knn(bool mixed, const KnnArgs &a, unsigned nb, hipStream_t s)
{
    if (mixed)
        hipLaunchKernelGGL((k_knn<KCAP, true>), dim3(nb), dim3(64), 0, s, a, a.nrec, a.cbox,
                           a.fbox, a.spos);
    else
        hipLaunchKernelGGL((k_knn<KCAP, false>), dim3(nb), dim3(64), 0, s, a, a.nrec, a.cbox,
                           a.fbox, a.spos);
}

static void dispatch_knn(int K, bool mixed, const KnnArgs &a, unsigned nb, hipStream_t s)
{
    // the mixed kernel is compiled without the explicit-query-list path
    if (mixed && a.qlist) throw std::logic_error("the mixed kNN kernel takes no query list");
    if (K <= 16) launch_knn<16>(mixed, a, nb, s);
    else if (K <= 32) launch_knn<32>(mixed, a, nb, s);
    else if (K <= 64) launch_knn<64>(mixed, a, nb, s);
    else launch_knn<128>(mixed, a, nb, s);
}

// The mixed kNN for k <= 64: the selection kernel on a persistent grid of as many waves as
// the device holds at once (each with its log region), taking the nw waves of queries from
// per-XCD counters.
template <int KCAP>
static void launch_sel(const KnnArgs &a, unsigned nw, int cus, DevBuf &klog, DevBuf &wctr,
                       hipStream_t s)
{
    int per_cu = 0;
    AHOP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_knn_sel<KCAP>, WAVE, 0));
    const unsigned g = std::max(1u, std::min(nw, (unsigned)(std::max(per_cu, 1) * cus)));
    uint2 *lg = static_cast<uint2 *>(klog.ensure((size_t)g * SEL_CAPL * WAVE * sizeof(uint2)));
    int *ctr = static_cast<int *>(wctr.ensure(8 * sizeof(int)));
    AHOP_HIP(hipMemsetAsync(ctr, 0, 8 * sizeof(int), s));
    hipLaunchKernelGGL((k_knn_sel<KCAP>), dim3(g), dim3(WAVE), 0, s, a, a.nrec, a.cbox, a.fbox,
                       a.spos, lg, ctr);
}

void Engine::launch_knn_sel(const KnnArgs &a, unsigned nw, hipStream_t s)
{
    if (sel_cus_ == 0) {
        int dev = 0;
        AHOP_HIP(hipGetDevice(&dev));
        AHOP_HIP(hipDeviceGetAttribute(&sel_cus_, hipDeviceAttributeMultiprocessorCount, dev));
    }
    if (a.K <= 16) launch_sel<16>(a, nw, sel_cus_, b_klog, b_wctr, s);
    else if (a.K <= 32) launch_sel<32>(a, nw, sel_cus_, b_klog, b_wctr, s);
    else launch_sel<64>(a, nw, sel_cus_, b_klog, b_wctr, s);
}

void Engine::knn_density(const Params &p)
{
    hipStream_t s = stream_;
    const int64_t n = t_.n;
    double *rho_s = as<double>(b_rho_s, n);
    int32_t *nbr = as<int32_t>(b_nbr, n * p.nhop);
    int32_t *fb = as<int32_t>(b_fallback, n);
    // [0] bad, [1] overflow, [2] fb count, [3] changed, [5] refit list, [6] refit stalled
    int *misc = (int *)b_misc.p;
    unsigned long long *cnt = as<unsigned long long>(b_counter, 4);
    AHOP_HIP(hipMemsetAsync(cnt, 0, 4 * sizeof(unsigned long long), s));
    KnnArgs a{};
    a.spos = t_.spos;
    a.first = t_.first;
    a.last = t_.last;
    a.left = t_.left;
    a.right = t_.right;
    a.box = t_.box;
    a.fbox = t_.fbox;
    a.nrec = t_.nrec;
    a.cbox = t_.cbox;
    a.n = (int)n;
    a.K = p.nvoisins;
    a.nhop = p.nhop;
    a.bucket = p.bucket;
    a.Lx = p.xlong;
    a.Ly = p.ylong;
    a.Lz = p.zlong;
    a.xyzl = p.xlong * p.ylong * p.zlong;
    a.rho_thr = p.rho_threshold;
    a.qlist = nullptr;
    a.nq = (int)n;
    a.qper = WAVE;
    a.rho_s = rho_s;
    a.nbr = nbr;
    a.fb_list = fb;
    a.fb_count = misc + 2;
    a.nact = cnt;
    a.overflow = misc + 1;
    unsigned long long *stats = as<unsigned long long>(b_counter, 24) + 8;
    AHOP_HIP(hipMemsetAsync(stats, 0, 16 * sizeof(unsigned long long), s));
    a.stats = stats;
    unsigned nb = (unsigned)((n + 63) / 64);
    a.nblocks = (int)nb;
    if (p.mixed && p.nvoisins <= 64) launch_knn_sel(a, nb, s);
    else dispatch_knn(p.nvoisins, p.mixed != 0, a, nb, s);
    AHOP_HIP(hipEventRecord(ev_[3], s));
    AHOP_HIP(hipMemcpyAsync(res_.knn_stats, stats, KNN_STAT_N * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, s));
    int host_misc[8];
    AHOP_HIP(hipMemcpyAsync(host_misc, misc, 8 * sizeof(int), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    if (host_misc[6] != 0) throw std::runtime_error("device tree: bounding-box refit stalled");
    if (host_misc[1] != 0) throw std::runtime_error("device tree deeper than the traversal stack");
    res_.nfallback = host_misc[2];
    if (host_misc[2] > 0) {
        // re-run the uncertified queries in FP64, in sorted-slot order so that each wave's
        // queries are as close together as the list allows
        const int nfb = host_misc[2];
        int32_t *fb_sorted = as<int32_t>(b_fb_sorted, nfb);
        size_t tb = 0;
        AHOP_HIP(rocprim::radix_sort_keys(nullptr, tb, fb, fb_sorted, (size_t)nfb, 0, 32, s));
        AHOP_HIP(rocprim::radix_sort_keys(b_temp.ensure(tb), tb, fb, fb_sorted, (size_t)nfb, 0, 32, s));
        KnnArgs b = a;
        b.qlist = fb_sorted;
        b.nq = nfb;
        // a short list is spread out over the box: one query per wave, so that no wave
        // walks the union of 64 unrelated search spheres
        b.qper = nfb <= 16384 ? 1 : WAVE;
        unsigned nb2 = (unsigned)((b.nq + b.qper - 1) / b.qper);
        b.nblocks = (int)nb2;
        b.stats = nullptr;
        if (p.nvoisins <= WAVE && n > p.nvoisins) {
            // one query per wave, the lanes sharing its search (4 041 queries at the 16.8M box:
            // 2.35 ms with the per-lane heap kernel, one lane busy per wave)
            hipLaunchKernelGGL(k_knn_one, dim3((unsigned)nfb), dim3(WAVE), 0, s, b);
        } else {
            dispatch_knn(p.nvoisins, false, b, nb2, s);
        }
    }
    AHOP_HIP(hipEventRecord(ev_[4], s));
}

void Engine::hop_groups(const Params &p)
{
    hipStream_t s = stream_;
    const int64_t n = t_.n;
    double *rho_s = (double *)b_rho_s.p;
    int32_t *nbr = (int32_t *)b_nbr.p;
    int32_t *hop_s = as<int32_t>(b_hop_s, n);
    int32_t *root = as<int32_t>(b_roots, n);
    int32_t *tmp = as<int32_t>(b_gid_s, n);
    int *misc = (int *)b_misc.p;
    hipLaunchKernelGGL(k_count_active, dim3(std::min(nblk(n), 2048u)), dim3(256), 0, s, rho_s, n,
                       p.rho_threshold, (unsigned long long *)b_counter.p);
    hipLaunchKernelGGL(k_hop, dim3(nblk(n)), dim3(256), 0, s, rho_s, nbr, t_.perm, n, p.nhop,
                       p.rho_threshold, hop_s);
    AHOP_HIP(hipMemcpyAsync(root, hop_s, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    // pointer doubling until every active particle points at its root: rounds in batches (4,
    // then 2 at a time), the host looking only at whether the batch's last round changed
    // anything (a round that changes nothing leaves every pointer at its root)
    for (int round = 0; round < 64;) {
        const int batch = round == 0 ? 4 : 2;
        for (int r = 0; r < batch; ++r, ++round) {
            if (r == batch - 1) AHOP_HIP(hipMemsetAsync(misc + 3, 0, sizeof(int), s));
            hipLaunchKernelGGL(k_jump, dim3(nblk(n)), dim3(256), 0, s, root, tmp, n, misc + 3);
            std::swap(root, tmp);
        }
        int changed = 0;
        AHOP_HIP(hipMemcpyAsync(&changed, misc + 3, sizeof(int), hipMemcpyDeviceToHost, s));
        AHOP_HIP(hipStreamSynchronize(s));
        if (!changed) break;
    }
    int32_t *flag_o = as<int32_t>(b_flag_o, n);
    int32_t *rank_o = as<int32_t>(b_rank_o, n);
    hipLaunchKernelGGL(k_root_flags, dim3(nblk(n)), dim3(256), 0, s, hop_s, t_.perm, n, flag_o);
    size_t tb = 0;
    AHOP_HIP(rocprim::exclusive_scan(nullptr, tb, flag_o, rank_o, 0, (size_t)n,
                                     rocprim::plus<int32_t>(), s));
    void *temp = b_temp.ensure(tb);
    AHOP_HIP(rocprim::exclusive_scan(temp, tb, flag_o, rank_o, 0, (size_t)n,
                                     rocprim::plus<int32_t>(), s));
    int32_t last2[2];
    AHOP_HIP(hipMemcpyAsync(&last2[0], rank_o + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipMemcpyAsync(&last2[1], flag_o + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    unsigned long long nact = 0;
    AHOP_HIP(hipMemcpyAsync(&nact, b_counter.p, sizeof(nact), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    res_.ngroups = last2[0] + last2[1];
    res_.nactive = (int64_t)nact;
    int32_t *gid_s = as<int32_t>(b_gid_s, n);  // tmp no longer needed (root holds the result)
    if (gid_s == root) {
        // root lives in b_gid_s after an odd number of swaps: move it out first
        AHOP_HIP(hipMemcpyAsync(as<int32_t>(b_keep, n), root, n * sizeof(int32_t),
                                hipMemcpyDeviceToDevice, s));
        root = (int32_t *)b_keep.p;
    }
    res_.density = as<double>(b_density, n);
    res_.gid = as<int32_t>(b_gid, n);
    res_.hop = as<int32_t>(b_hop, n);
    res_.densityg = as<double>(b_densityg, res_.ngroups);
    res_.nbr = nbr;
    hipLaunchKernelGGL(k_groups, dim3(nblk(n)), dim3(256), 0, s, root, hop_s, t_.perm, rank_o,
                       rho_s, n, gid_s, res_.gid, res_.hop, res_.density, res_.densityg);
    AHOP_HIP(hipEventRecord(ev_[5], s));
}

void Engine::saddles(const Params &p)
{
    hipStream_t s = stream_;
    const int64_t n = t_.n;
    const int G = res_.ngroups;
    int32_t *gid_s = (int32_t *)b_gid_s.p;
    double *rho_s = (double *)b_rho_s.p;
    int32_t *nbr = (int32_t *)b_nbr.p;
    int32_t *cnt = as<int32_t>(b_emit_cnt, n);
    int32_t *off = as<int32_t>(b_emit_off, n);
    hipLaunchKernelGGL(k_emit_count, dim3(nblk(n)), dim3(256), 0, s, gid_s, nbr, n, p.nhop, cnt);
    size_t tb = 0;
    AHOP_HIP(rocprim::exclusive_scan(nullptr, tb, cnt, off, 0, (size_t)n, rocprim::plus<int32_t>(), s));
    AHOP_HIP(rocprim::exclusive_scan(b_temp.ensure(tb), tb, cnt, off, 0, (size_t)n,
                                     rocprim::plus<int32_t>(), s));
    int32_t tail[2];
    AHOP_HIP(hipMemcpyAsync(&tail[0], off + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipMemcpyAsync(&tail[1], cnt + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    const int64_t ne = (int64_t)tail[0] + tail[1];
    res_.sad_offset = as<int32_t>(b_sad_off, G + 1);
    if (ne == 0 || G == 0) {
        AHOP_HIP(hipMemsetAsync(res_.sad_offset, 0, (G + 1) * sizeof(int32_t), s));
        res_.sad_isad = as<int32_t>(b_sad_isad, 1);
        res_.sad_rho = as<double>(b_sad_rho, 1);
        res_.nsaddle = 0;
        AHOP_HIP(hipEventRecord(ev_[6], s));
        return;
    }
    uint64_t *ekey = as<uint64_t>(b_ekey, ne);
    double *eval = as<double>(b_eval, ne);
    uint64_t *ekey2 = as<uint64_t>(b_ekey2, ne);
    double *eval2 = as<double>(b_eval2, ne);
    hipLaunchKernelGGL(k_emit_fill, dim3(nblk(n)), dim3(256), 0, s, gid_s, nbr, rho_s, off, n,
                       p.nhop, ekey, eval);
    unsigned gbits = 1;
    while ((1ull << gbits) <= (unsigned long long)G) ++gbits;
    unsigned end_bit = 32 + gbits;
    tb = 0;
    AHOP_HIP(rocprim::radix_sort_pairs(nullptr, tb, ekey, ekey2, eval, eval2, (size_t)ne, 0,
                                       end_bit, s));
    AHOP_HIP(rocprim::radix_sort_pairs(b_temp.ensure(tb), tb, ekey, ekey2, eval, eval2,
                                       (size_t)ne, 0, end_bit, s));
    uint64_t *ukey = as<uint64_t>(b_ukey, ne);
    double *uval = as<double>(b_uval, ne);
    int32_t *nuniq = (int32_t *)b_misc.p + 8;
    tb = 0;
    MaxOp maxop;
    AHOP_HIP(rocprim::reduce_by_key(nullptr, tb, ekey2, eval2, (size_t)ne, ukey, uval, nuniq,
                                    maxop, rocprim::equal_to<uint64_t>(), s));
    AHOP_HIP(rocprim::reduce_by_key(b_temp.ensure(tb), tb, ekey2, eval2, (size_t)ne, ukey, uval,
                                    nuniq, maxop, rocprim::equal_to<uint64_t>(), s));
    int32_t *keep = as<int32_t>(b_keep, ne);
    int32_t *koff = as<int32_t>(b_keep_off, ne);
    double *sval = as<double>(b_sval, ne);
    hipLaunchKernelGGL(k_symmetrize, dim3(nblk(ne)), dim3(256), 0, s, ukey, uval, nuniq, keep,
                       sval);
    int32_t hnu = 0;
    AHOP_HIP(hipMemcpyAsync(&hnu, nuniq, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    tb = 0;
    AHOP_HIP(rocprim::exclusive_scan(nullptr, tb, keep, koff, 0, (size_t)hnu, rocprim::plus<int32_t>(), s));
    AHOP_HIP(rocprim::exclusive_scan(b_temp.ensure(tb), tb, keep, koff, 0, (size_t)hnu,
                                     rocprim::plus<int32_t>(), s));
    int32_t kt[2];
    AHOP_HIP(hipMemcpyAsync(&kt[0], koff + hnu - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipMemcpyAsync(&kt[1], keep + hnu - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    const int64_t nk = (int64_t)kt[0] + kt[1];
    uint32_t *g1 = (uint32_t *)as<int32_t>(b_sad_cnt, nk);
    int32_t *g2 = as<int32_t>(b_emit_cnt, nk);       // emission counts are no longer needed
    double *sv = as<double>(b_eval, nk);
    hipLaunchKernelGGL(k_compact, dim3(nblk(hnu)), dim3(256), 0, s, ukey, sval, keep, koff, nuniq,
                       g1, g2, sv);
    hipLaunchKernelGGL(k_seg_offsets, dim3(nblk(G + 1)), dim3(256), 0, s, g1, nk, G,
                       res_.sad_offset);
    res_.sad_isad = as<int32_t>(b_sad_isad, nk);
    res_.sad_rho = as<double>(b_sad_rho, nk);
    tb = 0;
    AHOP_HIP(rocprim::segmented_radix_sort_pairs_desc(nullptr, tb, sv, res_.sad_rho, g2,
                                                      res_.sad_isad, (unsigned)nk, (unsigned)G,
                                                      res_.sad_offset, res_.sad_offset + 1, 0,
                                                      64, s));
    AHOP_HIP(rocprim::segmented_radix_sort_pairs_desc(b_temp.ensure(tb), tb, sv, res_.sad_rho, g2,
                                                      res_.sad_isad, (unsigned)nk, (unsigned)G,
                                                      res_.sad_offset, res_.sad_offset + 1, 0,
                                                      64, s));
    res_.nsaddle = nk;
    AHOP_HIP(hipEventRecord(ev_[6], s));
}

// sorted-slot results -> original order: density and the full neighbour rows (original ids)
__global__ void k_lists_to_original(const int32_t *__restrict__ nbr_s,
                                    const int32_t *__restrict__ perm,
                                    const double *__restrict__ rho_s, int64_t n, int K,
                                    int32_t *__restrict__ nbr_o, double *__restrict__ rho_o)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t o = perm[s];
    rho_o[o] = rho_s[s];
    for (int t = 0; t < K; ++t) nbr_o[o * K + t] = perm[nbr_s[s * K + t]];
}

static void validate(int64_t n, const Params &p)
{
    if (n < 2) throw std::runtime_error("need at least 2 particles");
    if (n >= ((int64_t)1 << 30)) throw std::runtime_error("particle count must be < 2^30");
    if (p.nvoisins < 1 || p.nvoisins >= n) throw std::runtime_error("need 1 <= nvoisins < npart");
    if (p.nvoisins > 128) throw std::runtime_error("nvoisins > 128 is not supported");
    if (p.nhop < 0 || p.nhop > p.nvoisins) throw std::runtime_error("need 0 <= nhop <= nvoisins");
    if (p.bucket < 1 || p.bucket > 64) throw std::runtime_error("bucket must be in [1, 64]");
    if (!(p.xlong > 0 && p.ylong > 0 && p.zlong > 0)) throw std::runtime_error("box sides must be > 0");
}

void Engine::knn_lists(const double *d_pos, const double *d_mass, int64_t n, const Params &p,
                       double *d_rho, int32_t *d_nbr, int32_t *d_order)
{
    validate(n, p);
    Params q = p;
    q.nhop = p.nvoisins;   // the kNN kernel stores all k ids per particle
    times_ = StageTimes{};
    AHOP_HIP(hipEventRecord(ev_[0], stream_));
    build_tree(d_pos, d_mass, n, q);
    knn_density(q);
    hipLaunchKernelGGL(k_lists_to_original, dim3(nblk(n)), dim3(256), 0, stream_,
                       (const int32_t *)b_nbr.p, t_.perm, (const double *)b_rho_s.p, n, q.nvoisins,
                       d_nbr, d_rho);
    if (d_order)
        AHOP_HIP(hipMemcpyAsync(d_order, t_.perm, sizeof(int32_t) * n, hipMemcpyDeviceToDevice,
                                stream_));
    AHOP_HIP(hipEventRecord(ev_[5], stream_));
    AHOP_HIP(hipStreamSynchronize(stream_));
    AHOP_HIP(hipEventElapsedTime(&times_.sort, ev_[0], ev_[1]));
    AHOP_HIP(hipEventElapsedTime(&times_.tree, ev_[1], ev_[2]));
    AHOP_HIP(hipEventElapsedTime(&times_.knn, ev_[2], ev_[3]));
    AHOP_HIP(hipEventElapsedTime(&times_.fallback, ev_[3], ev_[4]));
    AHOP_HIP(hipEventElapsedTime(&times_.total, ev_[0], ev_[5]));
    res_.nactive = 0;
    res_.ngroups = 0;
}

void Engine::run(const double *d_pos, const double *d_mass, int64_t n, const Params &p, bool saddle)
{
    validate(n, p);
    times_ = StageTimes{};
    AHOP_HIP(hipEventRecord(ev_[0], stream_));
    build_tree(d_pos, d_mass, n, p);
    knn_density(p);
    hop_groups(p);
    if (saddle) saddles(p);
    else AHOP_HIP(hipEventRecord(ev_[6], stream_));
    AHOP_HIP(hipStreamSynchronize(stream_));
    AHOP_HIP(hipEventElapsedTime(&times_.sort, ev_[0], ev_[1]));
    AHOP_HIP(hipEventElapsedTime(&times_.tree, ev_[1], ev_[2]));
    AHOP_HIP(hipEventElapsedTime(&times_.knn, ev_[2], ev_[3]));
    AHOP_HIP(hipEventElapsedTime(&times_.fallback, ev_[3], ev_[4]));
    AHOP_HIP(hipEventElapsedTime(&times_.hop, ev_[4], ev_[5]));
    AHOP_HIP(hipEventElapsedTime(&times_.saddle, ev_[5], ev_[6]));
    AHOP_HIP(hipEventElapsedTime(&times_.total, ev_[0], ev_[6]));
}

}  // namespace ahop
