// knn_sel.h -- the mixed-precision kNN + SPH density kernel for k <= 64 (included by
// ahop_kernels.hip after knn_wave.h, whose box tests, FP32 key error bound and heap layout
// constants it shares).
//
// As in k_knn, one wave serves 64 consecutive Morton-sorted queries, one per lane, walks the
// radix tree with them together, stages leaf buckets in LDS and defers insertions to
// lane-serial flushes.  What differs is what a lane keeps while it searches:
//  * only the FP32 keys of its k + 1 nearest candidates so far, in a max-heap whose root is a
//    register and whose two levels below (8 entries, then 8 under each) live in LDS at 4 B
//    per entry, in the 16 KB that later hold the sorted neighbour slots.  The wave's LDS is
//    18 KB instead of 30 KB, so 8 waves fit per CU instead of 5 (the registers allow 8
//    too), and a replacement moves keys only, with one LDS store per level (the first
//    level in registers needed 8 selects to write one entry back: C3 kernel 50.2 ->
//    48.4 ms).  Each key carries its place among its siblings in its low 3 mantissa bits,
//    so one integer max over a group finds the largest and where it is;
//  * every (key, sorted slot) it puts into the heap during the traversal is appended to a
//    per-lane log in global memory (one region per resident wave: the grid is persistent and
//    takes waves of queries from per-XCD counters); the seeding's candidates all lie in the
//    seed window, whose keys the collection computes again.  A key is evicted only as the
//    heap's maximum, and the maximum never grows, so after the search the window's and the
//    logged entries whose key is below the final root are exactly the heap's entries below
//    it: the lane's k nearest when they number k.
// The k survivors are then put in FP32-key order in registers (Batcher's odd-even merge
// network, ties by slot), their exact FP64 distances are checked to be in the same order
// (FP32 near-ties can swap two; such a query is re-run on the FP64 path, as are those whose
// log overflowed), and the density is summed in that order, as the reference does
// (omp_compute_density f90:2816-2838).
#pragma once
#include <utility>

// log entries per lane: the traversal's insertions (at the 16.8M-particle box at most 575;
// 2 of its 16.8M queries overflow 512 and are re-run on the FP64 path)
constexpr int SEL_CAPL = 512;
// slots of the wave's Morton seed window beyond the k/2 on either side of its queries (the
// heap kernels take 16): 32 gives the lanes at a wave's ends better seeds (at the
// 16.8M-particle box, 56.5 -> 55.9 ms, largest log 2 315 -> 640 entries, log overflows
// 938 -> 6; 8 took 62.7 and 24 61.1 against 61.5 ms with the untagged heap)
#ifndef SEL_SEED_PAD
#define SEL_SEED_PAD 32
#endif
// neighbours the density gathers at a time
#ifndef SEL_DCH
#define SEL_DCH 8
#endif
// log entries the collection reads ahead per lane
#ifndef SEL_P2U
#define SEL_P2U 8
#endif

// The heap holds keys as int bit patterns of the FP32 keys (non-negative floats and +-inf
// order as signed ints) with their low 3 mantissa bits replaced by a tag: an entry's position
// in its group of LC = 8 siblings.  One max over a group then gives both the largest key and
// where it sits (no compare-and-select chain for the index).  The heap orders the keys
// quantised to 8 ulps (q(key) = bits & ~7), its root is a quantised key, and every
// comparison with a candidate is on quantised values; the root still only decreases, so
// every candidate ever dropped (rejected or evicted) has a full key at or above the final
// root, and the certificate on full keys is unchanged.
constexpr int KQ_MASK = ~7;

__device__ __forceinline__ int kq(float k) { return __float_as_int(k) & KQ_MASK; }

// The bucket scan only picks which staged candidates a lane will look at again: a picked
// candidate's key is recomputed as above (SEL_KEY) and compared with the lane's (k+1)-th key
// tau before it can enter the heap.  The scan therefore tests the expanded form
//   e = fma(cx, -2qx, fma(cy, -2qy, fma(cz, -2qz, w))),  w = |c|^2 staged with the bucket,
// against the threshold below, which is never smaller than needed: whenever the key d_s
// (SEL_KEY, on the same FP32 values c and q) is below tau, e is below it.  With u = 2^-24 and D
// = |c - q|^2 = |c|^2 - 2 c.q + |q|^2 exactly:
//  * d_s >= (1 - u)^5 D (two roundings of each difference squared, three of the sums), so
//    d_s < tau gives D < tau (1 + 5.01u) and |c| <= |q| + sqrt(tau)(1 + 2.6u) =: S + r;
//  * |w - |c|^2| <= 3.01u |c|^2, and the three fmas add at most 3u (|w| + 2|c||q|), so
//    e <= D - |q|^2 + 6.02u (S + r)(S + 2r);
//  * the threshold (tau - Qn) + M, with Qn = |q|^2 (1 +- 3.01u) computed like w, loses at
//    most 3.01u |q|^2 + 2.01u (tau + |q|^2) to roundings.
// Together, e < threshold holds when M (1 - u) >= 14.3u (S + 2r)^2; M = 32u (sqrt(tau) +
// 2 sqrt(Qn))^2 clears that (the tiny absolute term guards against underflow).  Lanes
// without a search (tau < 0) get -inf: they accept nothing; tau = +inf accepts everything.
// The scan costs 5 VALU per candidate (3 fmas, the subtraction, and v_alignbit shifting the
// difference's sign bit into the lane's mask) against 10 for the key, a compare and a select.
__device__ __forceinline__ float sel_thr(float tau, float Qn, float rqn)
{
    if (!(tau >= 0.0f)) return -INFINITY;
    const float s = sqrtf(tau) + 2.0f * rqn;
    return (tau - Qn) + (32.0f * U32 * (s * s) + 1e-30f);
}

__device__ __forceinline__ int max8(int a, int b, int c, int d, int e, int f, int g, int h)
{
    return max(max(max(a, b), max(c, d)), max(max(e, f), max(g, h)));
}

// The heap below its root (a register) lives in LDS, entry e >= 1 at slot e - 1 of the lane's
// column: its LC first-level entries (slots 0..LC-1), then LC children under each of them
// (first-level entry c, slot c, has its children at slots LC(c+1) ..  LC(c+1) + LC-1; those
// past the heap's KHCAP entries are absent).  Keys are quantised and tagged with their place
// among their siblings, so one integer max over a group gives the largest and its place.

// Sift quantised key dq down from first-level entry c (its old key is being replaced): the
// largest of its children moves up if it exceeds dq, and dq takes its place.
template <int KHCAP>
__device__ __forceinline__ void kt_sift_l1(int *hw, int lane, int dq, int c)
{
    const int c0 = LC * (c + 1);
    int vk = dq;
    if (c0 + 1 < KHCAP) {
        int *base = hw + c0 * WAVE + lane;
        int kt[LC];
#pragma unroll
        for (int t = 0; t < LC; ++t) kt[t] = base[t * WAVE];
        const int m = max8(kt[0], kt[1], kt[2], kt[3], kt[4], kt[5], kt[6], kt[7]);
        const int mq = m & KQ_MASK;
        if (mq > dq) {
            const int tc = m & 7;
            base[tc * WAVE] = dq | tc;
            vk = mq;
        }
    }
    hw[c * WAVE + lane] = vk | c;
}

// sift quantised key dq down from the root k0
template <int KHCAP>
__device__ __forceinline__ void kt_sift_root(int &k0, int *hw, int lane, int dq)
{
    static_assert(LC == 8, "groups of 8 siblings");
    int kt[LC];
#pragma unroll
    for (int t = 0; t < LC; ++t) kt[t] = hw[t * WAVE + lane];
    const int m = max8(kt[0], kt[1], kt[2], kt[3], kt[4], kt[5], kt[6], kt[7]);
    const int mq = m & KQ_MASK;
    if (mq > dq) {
        k0 = mq;
        kt_sift_l1<KHCAP>(hw, lane, dq, m & 7);
    } else {
        k0 = dq;
    }
}

// write quantised key q at heap entry e (e = 0: the root), tagged with its place among its
// siblings
__device__ __forceinline__ void kt_set(int &k0, int *hw, int lane, int e, int q)
{
    if (e == 0) k0 = q;
    else hw[(e - 1) * WAVE + lane] = q | ((e - 1) & (LC - 1));
}

// Batcher's odd-even merge sorting network on N (a power of two) entries, as a list of
// compare-exchange pairs (a[i] < b[i]) built at compile time
template <int N>
struct SortNet {
    int n;
    int a[N * 16], b[N * 16];
};
template <int N>
constexpr SortNet<N> make_sort_net()
{
    SortNet<N> s{};
    s.n = 0;
    for (int p = 1; p < N; p += p)
        for (int k = p; k > 0; k /= 2)
            for (int j = k % p; j + k < N; j += k + k)
                for (int i = 0; i < k && i < N - j - k; ++i)
                    if ((i + j) / (p + p) == (i + j + k) / (p + p)) {
                        s.a[s.n] = i + j;
                        s.b[s.n] = i + j + k;
                        ++s.n;
                    }
    return s;
}

template <int N>
constexpr SortNet<N> SORT_NET = make_sort_net<N>();

// spline_ref without branches: both pieces evaluated in the same arithmetic and the one that
// applies selected, so the value is spline_ref's bit for bit
__device__ __forceinline__ double spline_sel(double x)
{
    const double x2 = x * x;
    const double x3 = x2 * x;
    const double in = 1.0 - 1.5 * x2 + 0.75 * x3;
    const double y = 2.0 - x;
    const double y2 = y * y;
    const double y3 = y2 * y;
    const double out = 0.25 * y3;
    return x <= 1.0 ? in : (x <= 2.0 ? out : 0.0);
}

// compare-exchange of entries IA < IB of packed (FP32 key bits, slot) words: keys are
// non-negative, so their bits order as the keys do, and equal keys order by slot
template <int IA, int IB, int N>
__device__ __forceinline__ void net_ce(uint64_t (&kv)[N])
{
    const uint64_t a = kv[IA], b = kv[IB];
    const bool sw = b < a;
    kv[IA] = sw ? b : a;
    kv[IB] = sw ? a : b;
}

// the whole network, every index a compile-time constant, so that the array stays in
// registers
template <int N, int... C>
__device__ __forceinline__ void net_sort(uint64_t (&kv)[N], std::integer_sequence<int, C...>)
{
    (net_ce<SORT_NET<N>.a[C], SORT_NET<N>.b[C]>(kv), ...);
}

template <int KCAP>
__global__ __launch_bounds__(64, 2) void k_knn_sel(KnnArgs A, const int4 *__restrict__ nrec,
                                                const float4 *__restrict__ ncbox,
                                                const float4 *__restrict__ nfbox,
                                                const double4 *__restrict__ spos,
                                                uint2 *__restrict__ klog, int *__restrict__ wctr)
{
    constexpr int KHCAP = KCAP + 1;
    static_assert(KCAP <= 64, "the selection kernel serves k <= 64");
    static_assert(KHCAP > LC + 1 && (KHCAP - 1 - LC) % LC == 0,
                  "every first-level entry has 0 or LC children");
    constexpr int NL = KHCAP - 1;   // the heap's LDS entries
    // the heap's LDS entries; after the search, the lane's sorted neighbour slots
    __shared__ int hw[(NL > KCAP ? NL : KCAP) * WAVE];
    __shared__ __align__(16) float sx[2 * WAVE];
    __shared__ __align__(16) float sy[2 * WAVE];
    __shared__ __align__(16) float sz[2 * WAVE];
    __shared__ __align__(16) float sw[2 * WAVE];   // |c|^2 of the staged candidates
    __shared__ int32_t stack[STACK];

    const int lane = threadIdx.x;
    const double Lx = A.Lx, Ly = A.Ly, Lz = A.Lz;
    const float Lxf = (float)Lx, Lyf = (float)Ly, Lzf = (float)Lz;
    const float mLx = 4.0f * U32 * Lxf, mLy = 4.0f * U32 * Lyf, mLz = 4.0f * U32 * Lzf;
    const float Lmin = fminf(Lxf, fminf(Lyf, Lzf));
    const bool root_is_bucket = (A.n <= A.bucket);
    // The traversal stops at nodes of at most twice the tree's bucket size and stages them 64
    // slots at a time (C3: 106.5 node visits per wave instead of 147.5, kernel about 1 % faster;
    // 3 and 4 times took longer).  Their node records exist: the tree writes the records of
    // the children of every node above its bucket size.
    const int vbucket = 2 * A.bucket;
    // this block's log: SEL_CAPL consecutive entries per lane, so that a lane's appends fill
    // whole cache lines (lane-interleaved entries left lines partly written: 55 GB written
    // per launch at the 16.8M-particle box, kernel 65.7 against 63.2 ms)
    const int64_t lg_off = ((int64_t)blockIdx.x * WAVE + lane) * SEL_CAPL;
    // waves of queries: XCD x (blocks are dispatched round-robin over the 8 XCDs) first takes
    // the contiguous range x of them, so that neighbouring waves share an L2, then helps the
    // others
    const int nw = A.nblocks;
    const int R = (nw + 7) / 8;
    const int x0 = blockIdx.x & 7;
    unsigned long long st_nodes = 0, st_buckets = 0, st_cand = 0, st_steps = 0, st_waves = 0;
    unsigned st_ins = 0, st_logmax = 0, st_ovf = 0, st_misorder = 0;

    for (int t8 = 0; t8 < 8;) {
        const int xr = (x0 + t8) & 7;
        int it = 0;
        if (lane == 0) it = atomicAdd(&wctr[xr], 1);
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
        const int lo = xr * R;
        const int hi = min(nw, lo + R);
        if (lo + it >= hi) {
            ++t8;
            continue;
        }
        const int wave = lo + it;
        ++st_waves;
        // k and the log base are re-read opaquely for every wave of queries: otherwise the
        // compiler hoists the 64 log-entry addresses and the e < k tests of the unrolled
        // loops out of this loop and keeps them live (spilled) across the whole search
        int K = A.K;
        int64_t lgo = lg_off;
        asm volatile("" : "+s"(K), "+v"(lgo));
        uint2 *const lg = klog + lgo;
        const int KH = K + 1;
        const int Wd = K / 2 + SEL_SEED_PAD;

        const int q = wave * WAVE + lane;
        const bool active = q < A.n;
        const double4 P = spos[active ? q : A.n - 1];
        // centre of the wave's queries: origin of the box-relative FP32 coordinates
        double cx, cy, cz;
        {
            const double BIG = 1e300;
            const double lx = wave_min(active ? P.x : BIG), hx = wave_max(active ? P.x : -BIG);
            const double ly = wave_min(active ? P.y : BIG), hy = wave_max(active ? P.y : -BIG);
            const double lz = wave_min(active ? P.z : BIG), hz = wave_max(active ? P.z : -BIG);
            cx = uniform_d(0.5 * (lx + hx));
            cy = uniform_d(0.5 * (ly + hy));
            cz = uniform_d(0.5 * (lz + hz));
        }
        const double rx = wrapd(P.x - cx, Lx), ry = wrapd(P.y - cy, Ly), rz = wrapd(P.z - cz, Lz);
        const float qx = (float)rx, qy = (float)ry, qz = (float)rz;
        // the scan's expanded form: -2q (exact) and |q|^2, sqrt(|q|^2) (sel_thr)
        const float m2qx = -2.0f * qx, m2qy = -2.0f * qy, m2qz = -2.0f * qz;
        const float Qn = __builtin_fmaf(qz, qz, __builtin_fmaf(qy, qy, qx * qx));
        const float rqn = sqrtf(Qn);
        const double hql = active ? fmax(fabs(rx), fmax(fabs(ry), fabs(rz))) : 0.0;
        // the error bounds of this lane's keys need only its own offset from the origin
        const float hQ = (float)hql * (1.0f + 4 * U32);
        // the FP32 frame is valid while search radius + query spread < L/2
        const bool wave_unsafe = wave_max(hql) >= 0.25 * fmin(Lx, fmin(Ly, Lz));
        const float pxf = (float)P.x, pyf = (float)P.y, pzf = (float)P.z;

        int k0 = kq(INFINITY);   // the heap's root
#pragma unroll
        for (int e = 1; e < KHCAP; ++e)
            hw[(e - 1) * WAVE + lane] = kq(e < KH ? INFINITY : -INFINITY) | ((e - 1) % LC);
        float tau = (active && !wave_unsafe) ? INFINITY : -1.0f;
        // the wave's seed window [qb - Wd, qb + 63 + Wd]: every lane seeds from all of it, and
        // the traversal skips it
        const int seed_lo = max(0, wave * WAVE - Wd);
        const int seed_hi = min(A.n - 1, wave * WAVE + WAVE - 1 + Wd);
        const int qmid = wave * WAVE + WAVE / 2;
        uint64_t m0 = 0, m1 = 0;   // accepted-candidate masks of the two staged buckets
        int f0 = 0, f1 = 0;        // first slot of each staged bucket
        int nbuf = 0;
        // heap entries written by the seeding (lanes without a search count as full)
        int filled = tau < 0.0f ? KH : 0;
        int nlog = 0;
        // The seeding logs nothing: all its candidates lie in the seed window, whose keys the
        // collection computes again, and the traversal skips the window's slots
        bool logon = false;

#define SEL_KEY(X, Y, Z, out)                                                                  \
    do {                                                                                       \
        const float dx_ = (X) - qx, dy_ = (Y) - qy, dz_ = (Z) - qz;                            \
        (out) = __builtin_fmaf(dz_, dz_, __builtin_fmaf(dy_, dy_, dx_ * dx_));                 \
    } while (0)

    // (past the log's end the last entry is overwritten: such a lane is re-run in FP64)
#define SEL_LOG(d, id)                                                                         \
    do {                                                                                       \
        if (logon) { /* wave-uniform */                                                        \
            lg[min(nlog, SEL_CAPL - 1)] = make_uint2(__float_as_uint(d), (uint32_t)(id));     \
            ++nlog;                                                                            \
        }                                                                                      \
    } while (0)

#define SEL_INSERT(d, id)                                                                      \
    do {                                                                                       \
        kt_sift_root<KHCAP>(k0, hw, lane, kq(d));                                              \
        SEL_LOG((d), (id));                                                                    \
        ++st_ins;                                                                              \
        tau = __int_as_float(k0);                                                            \
    } while (0)

#define SEL_HEAPIFY()                                                                          \
    do {                                                                                       \
        _Pragma("unroll") for (int c_ = LC - 1; c_ >= 0; --c_) if (LC * (c_ + 1) + 1 < KHCAP)  \
            kt_sift_l1<KHCAP>(hw, lane, hw[c_ * WAVE + lane] & KQ_MASK, c_);                   \
        kt_sift_root<KHCAP>(k0, hw, lane, k0);                                                 \
    } while (0)

    // this lane's next pending candidate (first staged bucket with a set bit, lowest bit
    // first), or +inf when it has none left (the stage index then stays in range)
#define SEL_PICK(dout, idout)                                                                  \
    do {                                                                                       \
        const bool s0_ = m0 != 0;                                                              \
        uint64_t mk_ = s0_ ? m0 : m1;                                                          \
        const bool any_ = mk_ != 0;                                                            \
        const int fsel_ = s0_ ? f0 : f1;                                                       \
        const int j_ = (int)__builtin_ctzll(mk_ | (1ull << 63)); /* 63 when empty */          \
        mk_ &= mk_ - 1;                                                                        \
        m0 = s0_ ? mk_ : m0;                                                                   \
        m1 = s0_ ? m1 : mk_;                                                                   \
        const int sidx_ = (s0_ ? 0 : WAVE) + j_;                                               \
        SEL_KEY(sx[sidx_], sy[sidx_], sz[sidx_], dout);                                        \
        (dout) = any_ ? (dout) : INFINITY;                                                     \
        (idout) = fsel_ + j_;                                                                  \
    } while (0)

    // flush while some lane's heap is not yet full: one candidate per lane per step, written
    // straight into the next free entry until the heap is full, then heapified once
#define SEL_FLUSH_FILL()                                                                       \
    do {                                                                                       \
        while (__any((m0 | m1) != 0)) {                                                        \
            ++st_steps;                                                                        \
            if ((m0 | m1) != 0) {                                                              \
                float d_;                                                                      \
                int id_;                                                                       \
                SEL_PICK(d_, id_);                                                             \
                if (filled < KH) {                                                             \
                    kt_set(k0, hw, lane, filled, kq(d_));                                      \
                    SEL_LOG(d_, id_);                                                          \
                    ++st_ins;                                                                  \
                    if (++filled == KH) {                                                      \
                        SEL_HEAPIFY();                                                         \
                        tau = __int_as_float(k0);                                            \
                    }                                                                          \
                } else if (d_ < tau) {                                                         \
                    SEL_INSERT(d_, id_);                                                       \
                }                                                                              \
            }                                                                                  \
        }                                                                                      \
        nbuf = 0;                                                                              \
    } while (0)

    // flush of full heaps: every step takes two pending candidates of each lane (their stage
    // reads in flight together) and inserts them smallest first while they beat the lane's
    // current (k+1)-th key
#define SEL_FLUSH_W()                                                                          \
    do {                                                                                       \
        while (__any((m0 | m1) != 0)) {                                                        \
            ++st_steps;                                                                        \
            float a0_, a1_;                                                                    \
            int i0_, i1_;                                                                      \
            SEL_PICK(a0_, i0_);                                                                \
            SEL_PICK(a1_, i1_);                                                                \
            if (a1_ < a0_) {                                                                   \
                const float ta_ = a0_;                                                         \
                const int ti_ = i0_;                                                           \
                a0_ = a1_;                                                                     \
                i0_ = i1_;                                                                     \
                a1_ = ta_;                                                                     \
                i1_ = ti_;                                                                     \
            }                                                                                  \
            while (__any(a0_ < tau)) {                                                         \
                if (a0_ < tau) SEL_INSERT(a0_, i0_);                                           \
                a0_ = a1_;                                                                     \
                i0_ = i1_;                                                                     \
                a1_ = INFINITY;                                                                \
            }                                                                                  \
        }                                                                                      \
        nbuf = 0;                                                                              \
    } while (0)

#define SEL_FLUSH_S()                                                                          \
    do {                                                                                       \
        if (__all(filled >= KH)) SEL_FLUSH_W();                                                \
        else SEL_FLUSH_FILL();                                                                 \
    } while (0)

    // scan the staged candidates j = 0..cnt-1 of buffer b and set bit j of mask_ where the
    // expanded-form test against thr passes (sel_thr).  Each 32-candidate half is walked
    // from its top group of 4 down, every candidate's sign bit shifted in at the bottom, so
    // that candidate j ends at bit j; bits of the top group past cnt are cleared at the end
#define SEL_SCAN(b, cnt, thr)                                                                  \
    do {                                                                                       \
        uint32_t hm_[2] = {0u, 0u};                                                            \
        _Pragma("unroll") for (int h_ = 0; h_ < 2; ++h_) {                                     \
            const int nh_ = min((cnt), 32 * h_ + 32) - 32 * h_;                                \
            for (int j4_ = 32 * h_ + ((nh_ + 3) & ~3) - 4; j4_ >= 32 * h_; j4_ -= 4) {         \
                const float4 X_ = *reinterpret_cast<const float4 *>(sx + (b) * WAVE + j4_);    \
                const float4 Y_ = *reinterpret_cast<const float4 *>(sy + (b) * WAVE + j4_);    \
                const float4 Z_ = *reinterpret_cast<const float4 *>(sz + (b) * WAVE + j4_);    \
                const float4 W_ = *reinterpret_cast<const float4 *>(sw + (b) * WAVE + j4_);    \
                const float xs_[4] = {X_.x, X_.y, X_.z, X_.w};                                 \
                const float ys_[4] = {Y_.x, Y_.y, Y_.z, Y_.w};                                 \
                const float zs_[4] = {Z_.x, Z_.y, Z_.z, Z_.w};                                 \
                const float ws_[4] = {W_.x, W_.y, W_.z, W_.w};                                 \
                _Pragma("unroll") for (int t_ = 3; t_ >= 0; --t_) {                            \
                    const float e_ = __builtin_fmaf(                                           \
                        xs_[t_], m2qx,                                                         \
                        __builtin_fmaf(ys_[t_], m2qy, __builtin_fmaf(zs_[t_], m2qz, ws_[t_]))); \
                    hm_[h_] = __builtin_amdgcn_alignbit(hm_[h_], __float_as_uint(e_ - (thr)),  \
                                                        31u);                                  \
                }                                                                              \
            }                                                                                  \
        }                                                                                      \
        mask_ = ((uint64_t)hm_[1] << 32) | hm_[0];                                             \
        if ((cnt) < WAVE) mask_ &= (1ull << (cnt)) - 1ull;                                     \
    } while (0)

    // stage sorted slots [f, f+cnt) (cnt <= 64) as FP32 coordinates relative to the wave
    // centre and record each lane's accepted subset: seeding accepts only the seed window,
    // the traversal everything outside it.  The seeding flushes a full stage before staging
    // the next bucket, the traversal right after filling it.
#define SEL_STAGE(fst, cnt, seeding)                                                           \
    do {                                                                                       \
        if ((seeding) && nbuf == 2) SEL_FLUSH_S();                                             \
        __syncthreads();                                                                       \
        if (lane < (cnt)) {                                                                    \
            const double4 c_ = spos[(fst) + lane];                                             \
            const float x_ = (float)wrapd(c_.x - cx, Lx);                                      \
            const float y_ = (float)wrapd(c_.y - cy, Ly);                                      \
            const float z_ = (float)wrapd(c_.z - cz, Lz);                                      \
            sx[nbuf * WAVE + lane] = x_;                                                       \
            sy[nbuf * WAVE + lane] = y_;                                                       \
            sz[nbuf * WAVE + lane] = z_;                                                       \
            sw[nbuf * WAVE + lane] = __builtin_fmaf(z_, z_, __builtin_fmaf(y_, y_, x_ * x_));  \
        }                                                                                      \
        __syncthreads();                                                                       \
        uint64_t mask_ = 0;                                                                    \
        const float thr_ = sel_thr(tau, Qn, rqn);                                              \
        SEL_SCAN(nbuf, cnt, thr_);                                                             \
        if ((seeding) || ((fst) <= seed_hi && (fst) + (cnt) - 1 >= seed_lo)) {                 \
            /* seeding keeps the seed window's slots, the traversal the others; neither  */    \
            /* the lane's own slot (buckets outside the window hold no query of the wave) */   \
            const uint64_t win_ = bits_in(seed_lo - (fst), seed_hi - (fst) + 1);               \
            mask_ &= (seeding) ? win_ : ~win_;                                                 \
            const int so_ = q - (fst);                                                         \
            mask_ &= (so_ >= 0 && so_ < WAVE) ? ~(1ull << so_) : ~0ull;                        \
        }                                                                                      \
        m0 = nbuf == 0 ? mask_ : m0;                                                           \
        m1 = nbuf == 1 ? mask_ : m1;                                                           \
        f0 = nbuf == 0 ? (fst) : f0;                                                           \
        f1 = nbuf == 1 ? (fst) : f1;                                                           \
        ++nbuf;                                                                                \
    } while (0)

        // ---- seed: the wave's seed window, staged bucket by bucket
        for (int f = seed_lo; f <= seed_hi; f += WAVE) {
            const int cnt = min(WAVE, seed_hi - f + 1);
            SEL_STAGE(f, cnt, true);
            // Two full buckets of the window are staged, and every lane with a search has
            // accepted all of them but its own slot: KH of them, centred on the lane's own
            // slot as far as the two buckets allow, go straight into its heap (its own slot,
            // if among them, with key +inf, which the first insertion evicts).
            if (nbuf == 2 && f == seed_lo + WAVE && cnt == WAVE) {
                const int off = q - seed_lo;
                const uint64_t self0 = off < WAVE ? 1ull << off : 0ull;
                const uint64_t self1 = (off >= WAVE && off < 2 * WAVE) ? 1ull << (off - WAVE) : 0ull;
                if (m0 == ~self0 && m1 == ~self1) {
                    const int rs = min(max(off - KH / 2, 0), 2 * WAVE - KH);
                    for (int e = 0; e < KH; ++e) {
                        const int r = rs + e;
                        float d;
                        SEL_KEY(sx[r], sy[r], sz[r], d);
                        d = r == off ? INFINITY : d;
                        if (e == 0) k0 = kq(d);   /* wave-uniform */
                        else hw[(e - 1) * WAVE + lane] = kq(d) | ((e - 1) % LC);
                    }
                    filled = KH;
                    st_ins += KH;
                    SEL_HEAPIFY();
                    tau = __int_as_float(k0);
                    m0 &= ~bits_in(rs, rs + KH);
                    m1 &= ~bits_in(rs - WAVE, rs + KH - WAVE);
                }
            }
        }
        SEL_FLUSH_S();
        if (filled < KH) {
            // a clipped window (array ends): the unfilled entries stay +inf
            SEL_HEAPIFY();
            filled = KH;
            tau = __int_as_float(k0);
        }

        // ---- wave-coherent traversal of the binary radix tree (as k_knn)
        logon = true;
        int sp = 1;
        stack[0] = 0;
        bool stack_overflow = false;
        bool top_fresh = false, top_need = false;
        while (sp > 0) {
            const int node = __builtin_amdgcn_readfirstlane(stack[--sp]);
            int4 rec;
            float4 cb0, cb1, cb2, cb3;
            if (node < 0) {
                rec = make_int4(0, 0, ~node, ~node);
                cb0 = cb1 = cb2 = cb3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            } else {
                rec = nrec[node];
                cb0 = ncbox[4 * node];
                cb1 = ncbox[4 * node + 1];
                cb2 = ncbox[4 * node + 2];
                cb3 = ncbox[4 * node + 3];
            }
            const int f = rec.z, l = rec.w;
            ++st_nodes;
            // the visit threshold on box distances, the same for this node and its children
            const float tn = tau + 2 * fp32_key_err(tau, hQ, Lmin);
            bool need = top_need;
            if (!top_fresh) {
                const float bd = box_dist2_lower(node, spos, nfbox, pxf, pyf, pzf, Lxf, Lyf,
                                                 Lzf, mLx, mLy, mLz);
                need = active && bd < tn;
            }
            top_fresh = false;
            if (!__any(need)) continue;
            if (l - f + 1 <= vbucket || (root_is_bucket && node == 0)) {
                // a bucket inside the seed window was seen by every lane already
                if (f >= seed_lo && l <= seed_hi) continue;
                const int cnt = l - f + 1;
                ++st_buckets;
                st_cand += cnt;
                for (int h = 0; h < cnt; h += WAVE) {
                    SEL_STAGE(f + h, min(cnt - h, WAVE), false);
                    if (nbuf == 2) SEL_FLUSH_W();
                }
                continue;
            }
            const float bd1 = box_dist2_at(cb0.x, cb0.y, cb0.z, cb0.w, cb1.x, cb1.y, pxf, pyf, pzf,
                                           Lxf, Lyf, Lzf, mLx, mLy, mLz);
            const float bd2 = box_dist2_at(cb2.x, cb2.y, cb2.z, cb2.w, cb3.x, cb3.y, pxf, pyf, pzf,
                                           Lxf, Lyf, Lzf, mLx, mLy, mLz);
            const bool n1 = active && bd1 < tn, n2 = active && bd2 < tn;
            top_fresh = true;
            const bool a1 = __any(n1), a2 = __any(n2);
            if (sp + 2 > STACK) {
                stack_overflow = true;   // reported after the loop
                break;
            }
            if (a1 && a2) {
                // first the child that more lanes need closely (box distance under half the
                // threshold), then the one more lanes need at all; on a tie, the nearer in
                // sorted-slot distance to the wave's queries
                int k1 = __popcll(__ballot(n1 && bd1 < 0.5f * tn));
                int k2 = __popcll(__ballot(n2 && bd2 < 0.5f * tn));
                if (k1 == k2) {
                    k1 = __popcll(__ballot(n1));
                    k2 = __popcll(__ballot(n2));
                }
                const int f1_ = __float_as_int(cb1.z), l1_ = __float_as_int(cb1.w);
                const int f2_ = __float_as_int(cb3.z), l2_ = __float_as_int(cb3.w);
                const int d1 = max(0, max(f1_ - qmid, qmid - l1_));
                const int d2 = max(0, max(f2_ - qmid, qmid - l2_));
                if (k1 > k2 || (k1 == k2 && d1 <= d2)) {
                    stack[sp++] = rec.y;
                    stack[sp++] = rec.x;
                    top_need = n1;
                } else {
                    stack[sp++] = rec.x;
                    stack[sp++] = rec.y;
                    top_need = n2;
                }
            } else if (a1) {
                stack[sp++] = rec.x;
                top_need = n1;
            } else if (a2) {
                stack[sp++] = rec.y;
                top_need = n2;
            } else {
                top_fresh = false;
            }
        }
        SEL_FLUSH_W();
        if (stack_overflow && lane == 0) atomicAdd(A.overflow, 1);
#undef SEL_STAGE
#undef SEL_SCAN
#undef SEL_FLUSH_S
#undef SEL_FLUSH_W
#undef SEL_FLUSH_FILL
#undef SEL_PICK
#undef SEL_HEAPIFY
#undef SEL_INSERT
#undef SEL_LOG

        // ---- collection: the logged entries below the final (k+1)-th key, compacted in
        // place to the front of the lane's log (entry e is read before any write reaches it)
        // the final root: the (k+1)-th smallest key, quantised (on the 8-ulp grid, so a key is
        // below it exactly when its quantised value is)
        const float T = __int_as_float(k0);
        const int nl = min(nlog, SEL_CAPL);
        st_logmax = max(st_logmax, (unsigned)nlog);
        st_ovf += (nlog > SEL_CAPL) ? 1u : 0u;
        int cnt = 0;
        int kmaxb = 0;   // the largest collected key's bits (keys are non-negative)
        for (int e = 0; __any(e < nl); e += SEL_P2U) {
            // a lane reads only its own entries (reading every lane up to the wave's longest
            // log: C3 kernel 55.9 -> 56.8 ms)
            uint2 v[SEL_P2U];
#pragma unroll
            for (int u = 0; u < SEL_P2U; ++u)
                v[u] = e + u < nl ? lg[e + u] : make_uint2(0x7f800000u, 0u);
#pragma unroll
            for (int u = 0; u < SEL_P2U; ++u) {
                if (__uint_as_float(v[u].x) < T) {
                    if (cnt < KCAP) lg[cnt] = v[u];
                    ++cnt;
                    kmaxb = max(kmaxb, (int)v[u].x);
                }
            }
        }
        // the seed window's entries below T, from its buckets staged again (every log entry
        // is read by now, so they may go anywhere): every window slot with a key below T went
        // into the heap during the seeding, since the (k+1)-th key never fell below T
        for (int wf = seed_lo; wf <= seed_hi; wf += 2 * WAVE) {
            const int wn = min(2 * WAVE, seed_hi - wf + 1);
            __syncthreads();
            for (int b = 0; b < 2; ++b) {
                const int r = b * WAVE + lane;
                if (r < wn) {
                    const double4 c_ = spos[wf + r];
                    sx[r] = (float)wrapd(c_.x - cx, Lx);
                    sy[r] = (float)wrapd(c_.y - cy, Ly);
                    sz[r] = (float)wrapd(c_.z - cz, Lz);
                }
            }
            __syncthreads();
            const int off = q - wf;
            for (int r = 0; r < wn; ++r) {
                float kk;
                SEL_KEY(sx[r], sy[r], sz[r], kk);
                if (r != off && kk < T) {
                    if (cnt < KCAP) lg[cnt] = make_uint2(__float_as_uint(kk), (uint32_t)(wf + r));
                    ++cnt;
                    kmaxb = max(kmaxb, __float_as_int(kk));
                }
            }
        }
#undef SEL_KEY
        // Certificate (as k_knn): the k kept are the true k nearest when the (k+1)-th key
        // clears the k-th by twice the error bound E of a key of that size and every particle
        // within that radius lies in the periodic image the wave-centred frame picks.  Ties
        // at the (k+1)-th key leave fewer than k below it: re-run.  An unfilled heap (final
        // root +inf) holds every other particle.
        bool ok = active && !wave_unsafe && nlog <= SEL_CAPL && cnt == K;
        if (ok && T != INFINITY) {
            const float E = fp32_key_err(T, hQ, Lmin);
            const float kmax = __int_as_float(kmaxb);
            ok = T - kmax > 2 * E && E < 0.01f * T && sqrtf(T * 1.0001f) + hQ < 0.4999f * Lmin;
        }
        if (active && !ok) A.fb_list[atomicAdd(A.fb_count, 1)] = q;
        if (ok) {
            // ---- the k kept in FP32-key order (ties by slot), sorted in registers; then their
            // exact FP64 distances (d2_ref, the reference's arithmetic) in that order: where
            // FP32 near-ties left two of them out of FP64 order, the query is re-run on the
            // FP64 path
            uint64_t kv[KCAP];
#pragma unroll
            for (int e = 0; e < KCAP; ++e) {
                // every log has more than KCAP entries: all are loaded, and those past k are
                // masked to all ones arithmetically (a select would become a branch around
                // each load, serialising them)
                const uint2 v = lg[e];
                kv[e] = (((uint64_t)v.x << 32) | v.y) | (0ull - (uint64_t)(e >= K));
            }
            net_sort<KCAP>(kv, std::make_integer_sequence<int, SORT_NET<KCAP>.n>{});
            // the sorted slots go to LDS (the heap is no longer needed), read back by the
            // runtime loops below
            uint32_t *const hid = reinterpret_cast<uint32_t *>(hw) + lane;
#pragma unroll
            for (int e = 0; e < KCAP; ++e) hid[e * WAVE] = (uint32_t)kv[e];
            // density in the reference's order: the k-th distance sets the smoothing length,
            // the first k-1 neighbours are summed in ascending distance
            const double4 Q = spos[q];
            const double4 ck = spos[hid[(K - 1) * WAVE]];
            const double d2k = d2_ref(Q.x, Q.y, Q.z, ck.x, ck.y, ck.z, Lx, Ly, Lz);
            const double r = sqrt(d2k) * 0.5;
            const double r3 = r * r * r;
            const double unsr = 1.0 / r;
            double contrib = 0.0;
            double prev = -1.0;
            bool sorted = true;
            // SEL_DCH neighbours gathered at a time; the terms past k-1 (on the query itself,
            // a valid slot) add +0.0, which leaves the sum's bits unchanged
            for (int i0 = 0; i0 < K - 1; i0 += SEL_DCH) {
                double4 c[SEL_DCH];
#pragma unroll
                for (int u = 0; u < SEL_DCH; ++u) {
                    const int i = i0 + u;
                    const uint32_t gi = hid[min(i, KCAP - 1) * WAVE];
                    c[u] = spos[i < K - 1 ? gi : (uint32_t)q];
                }
#pragma unroll
                for (int u = 0; u < SEL_DCH; ++u) {
                    const bool use = i0 + u < K - 1;
                    const double d = d2_ref(Q.x, Q.y, Q.z, c[u].x, c[u].y, c[u].z, Lx, Ly, Lz);
                    sorted = sorted && (d >= prev || !use);
                    prev = use ? d : prev;
                    const double term = c[u].w * spline_sel(sqrt(d) * unsr);
                    contrib = contrib + (use ? term : 0.0);
                }
            }
            sorted = sorted && d2k >= prev;
            if (sorted) {
                A.rho_s[q] = A.xyzl * (contrib + Q.w) / (PI_REF * r3);
                for (int t = 0; t < A.nhop; ++t)
                    A.nbr[(int64_t)q * A.nhop + t] = (int32_t)hid[t * WAVE];
            } else {
                ++st_misorder;
                A.fb_list[atomicAdd(A.fb_count, 1)] = q;
            }
        }
    }

    if (A.stats) {
        unsigned long long ins = st_ins;
        unsigned long long ovf = st_ovf;
        unsigned long long mis = st_misorder;
        unsigned lmax = st_logmax;
        for (int o = 32; o > 0; o >>= 1) {
            ins += __shfl_xor(ins, o);
            ovf += __shfl_xor(ovf, o);
            mis += __shfl_xor(mis, o);
            lmax = max(lmax, (unsigned)__shfl_xor((int)lmax, o));
        }
        if (lane == 0 && st_waves > 0) {
            atomicAdd(&A.stats[KNN_STAT_NODES], st_nodes);
            atomicAdd(&A.stats[KNN_STAT_BUCKETS], st_buckets);
            atomicAdd(&A.stats[KNN_STAT_CANDIDATES], st_cand);
            atomicAdd(&A.stats[KNN_STAT_INSERT_STEPS], st_steps);
            atomicAdd(&A.stats[KNN_STAT_INSERTS], ins);
            atomicAdd(&A.stats[KNN_STAT_WAVES], st_waves);
            atomicMax(&A.stats[KNN_STAT_LOG_MAX], (unsigned long long)lmax);
            atomicAdd(&A.stats[KNN_STAT_LOG_OVERFLOW], ovf);
            atomicAdd(&A.stats[KNN_STAT_MISORDER], mis);
        }
    }
}
