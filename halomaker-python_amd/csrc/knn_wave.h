// knn_wave.h -- the kNN + SPH density kernel (included by ahop_kernels.hip).
//
// One wave handles 64 consecutive Morton-sorted query particles, one query per lane.
// The wave walks the binary radix tree together (a node is entered when any lane needs
// it), stages the particles of each leaf bucket in LDS and every lane keeps the k (mixed:
// k+1) nearest candidates seen so far in a max-heap: mixed, the root and its 8 children (16
// above k = 64) in registers and their children in LDS; FP64, an 8-ary heap in LDS.
//
// The per-wave state is kept in plain local variables and the repeated steps are macros:
// a state struct (or closures capturing by reference) made the compiler spill it to
// scratch memory.
#pragma once


// conservative FP32 visit test of a node (c >= 0) or particle (c < 0) box: centres rounded
// to nearest, half-widths rounded up, and the margin mL covers the rounding of the query,
// the centre and the periodic fold, so the returned squared distance never exceeds the
// true minimum distance from the query to the box
__device__ __forceinline__ float box_dist2_at(float bx, float by, float bz, float hx, float hy,
                                              float hz, float px, float py, float pz, float Lx,
                                              float Ly, float Lz, float mLx, float mLy, float mLz)
{
    float tx = fabsf(px - bx), ty = fabsf(py - by), tz = fabsf(pz - bz);
    tx = fminf(tx, Lx - tx);
    ty = fminf(ty, Ly - ty);
    tz = fminf(tz, Lz - tz);
    const float dx = fmaxf(tx - hx - mLx, 0.0f);
    const float dy = fmaxf(ty - hy - mLy, 0.0f);
    const float dz = fmaxf(tz - hz - mLz, 0.0f);
    return (dx * dx + dy * dy + dz * dz) * (1.0f - 4 * U32);
}

__device__ __forceinline__ float box_dist2_lower(int c, const double4 *__restrict__ spos,
                                                 const float4 *__restrict__ fbox, float px,
                                                 float py, float pz, float Lx, float Ly,
                                                 float Lz, float mLx, float mLy, float mLz)
{
    if (c < 0) {
        const double4 p = spos[~c];
        return box_dist2_at((float)p.x, (float)p.y, (float)p.z, 0.0f, 0.0f, 0.0f, px, py, pz,
                            Lx, Ly, Lz, mLx, mLy, mLz);
    }
    const float4 b0 = fbox[2 * c], b1 = fbox[2 * c + 1];
    return box_dist2_at(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, px, py, pz, Lx, Ly, Lz, mLx, mLy, mLz);
}

// Bound on |d2f - d2| for a candidate at true squared distance ~d from a query, when both
// positions are rounded to FP32 relative to a common origin no farther than hQ from the
// query: per axis the FP32 difference is off by at most 2u(|dx| + hQ) (u = 2^-24), which
// with the rounding of the three squares and two sums gives
//   |d2f - d2| <= 7.01 u d + 6.94 u hQ sqrt(d) + 13 u^2 (sqrt(d) + hQ)^2 .
// The bound below rounds the constants up, evaluates at a slightly larger d, and adds a
// term for the FP64 rounding of the box-relative coordinates.
__device__ __forceinline__ float fp32_key_err(float d, float hQ, float L)
{
    const float s = sqrtf(d * 1.0001f) + hQ;
    return 8 * U32 * sqrtf(d * 1.0001f) * s + 16 * U32 * U32 * s * s + 1e-15f * L * s;
}

// a staged candidate: FP32 coordinates from the x, y, z arrays (mixed) or the FP64 double4
template <bool MIXED>
__device__ __forceinline__ auto stage_at(const float *sx, const float *sy, const float *sz,
                                         const double4 *st, int idx)
{
    if constexpr (MIXED) return make_float4(sx[idx], sy[idx], sz[idx], 0.0f);
    else return st[idx];
}

// the bits [lo, hi) of a 64-bit mask (clipped to 0..63)
__device__ __forceinline__ uint64_t bits_in(int lo, int hi)
{
    lo = max(lo, 0);
    hi = min(hi, 64);
    if (hi <= lo) return 0ull;
    const uint64_t top = hi >= 64 ? ~0ull : (1ull << hi) - 1ull;
    return top & ~((1ull << lo) - 1ull);
}

// The tree arrays and the sorted particles are also passed as restrict-qualified kernel
// arguments: the kernel never writes them, and saying so lets the wave-uniform node reads
// go through scalar loads.  A node visit reads the node's record (links, slot range) and its
// children's boxes and slot ranges from per-node arrays indexed by the node itself, so all of
// it comes in one round of scalar loads (with the children's boxes indexed by the children,
// the second round waited on the first: C3 kernel 104.3 -> 99.1 ms).  Two waves per SIMD caps the registers at 256 per lane, so that
// the 6 waves per CU the mixed kernel's LDS allows can run (without the cap it takes 258
// and runs 4 waves per CU).
template <int KCAP, bool MIXED>
__global__ __launch_bounds__(64, 2) void k_knn(KnnArgs A, const int4 *__restrict__ nrec,
                                            const float4 *__restrict__ ncbox,
                                            const float4 *__restrict__ nfbox,
                                            const double4 *__restrict__ spos)
{
    using KeyT = typename std::conditional<MIXED, float, double>::type;
    // the mixed heap keeps KX entries beyond the k nearest (the certificate needs the
    // (k+1)-th key) and holds KCAP + 1 entries: its LDS part (KCAP + 1 - RT entries) is
    // then a whole number of LC-child groups, so every register entry with children has
    // all LC of them allocated and insertions read them unchecked, the entries past the
    // heap size being kept at -inf.  The FP64 heap stays at KCAP entries (one more would
    // cost it a wave per CU of LDS) and checks child indices.
    constexpr int KX = 1;
    constexpr int KHCAP = MIXED ? KCAP + 1 : KCAP;
    constexpr int RC = heap_rc<KCAP>();
    constexpr int RT = RC + 1;
    constexpr bool SIFT_CHECKED = !MIXED;
    static_assert(KCAP % HEAP_D == 0, "heap capacity must be a multiple of the arity");
    // staged buckets whose insertions are deferred (the FP64 heap is twice as large, so its
    // stage may be cut down to fit one more wave per CU)
    constexpr int NBUF = MIXED ? AHOP_MIXED_NBUF : AHOP_FP64_NBUF;
    // heap storage: for the mixed kernel the root and its RC children in registers (rt) and
    // their children, packed (key, id) words, in LDS; key and id arrays for FP64
    static_assert(!MIXED || KHCAP <= RT + LC * RC, "the mixed heap has three levels");
    static_assert(!MIXED || (KHCAP - RT) % LC == 0, "every internal register entry has LC children");
    constexpr int NLDS = MIXED ? (KHCAP > RT ? KHCAP - RT : 1) : 1;
    __shared__ uint2 hp[NLDS * WAVE];
    __shared__ KeyT hk[MIXED ? 1 : KHCAP * WAVE];
    __shared__ int32_t hi[MIXED ? 1 : KHCAP * WAVE];
    RegTop<RC> rt;
    rt.k0 = -INFINITY;
    rt.i0 = 0u;
    rt.rk = -INFINITY;
    rt.ri = 0u;
#define HK(e) heap_key<MIXED, KeyT>(rt, hp, hk, lane, (e))
#define HI(e) heap_id<MIXED, KeyT>(rt, hp, hi, lane, (e))
#define HSET(e, k, i) heap_set<MIXED, KeyT>(rt, hp, hk, hi, lane, (e), (k), (i))
#define HSIFT(size, d, id, e, checked) \
    heap_sift_any<MIXED, checked, KeyT>(rt, hp, hk, hi, lane, (size), (d), (id), (e))
    // staged bucket coordinates: FP32 x, y, z arrays (mixed; three reads of four candidates
    // each serve four candidates) or FP64 double4 (FP64)
    __shared__ __align__(16) float sx[MIXED ? NBUF * WAVE : 4];
    __shared__ __align__(16) float sy[MIXED ? NBUF * WAVE : 4];
    __shared__ __align__(16) float sz[MIXED ? NBUF * WAVE : 4];
    __shared__ double4 stage[MIXED ? 1 : NBUF * WAVE];
    __shared__ int32_t stack[STACK];

    const int lane = threadIdx.x;
    const int wave = logical_block(A.nblocks);
    // explicit query lists are for the FP64 rerun only (the host checks it), so the mixed
    // kernel is compiled without that path
    const int32_t *const qlist = MIXED ? nullptr : A.qlist;
    // qper < WAVE only with an explicit query list (the seeding below assumes full waves
    // of consecutive slots otherwise)
    const int slot = wave * A.qper + lane;
    const bool active = lane < A.qper && slot < A.nq;
    const int q = active ? (qlist ? qlist[slot] : slot) : 0;
    const int K = A.K;
    const int KH = MIXED ? K + KX : K;
    const double4 P = spos[q];
    const double Lx = A.Lx, Ly = A.Ly, Lz = A.Lz;

    // centre of the wave's queries: origin of the box-relative FP32 coordinates
    double cx, cy, cz;
    {
        const double BIG = 1e300;
        const double lx = wave_min(active ? P.x : BIG), hx = wave_max(active ? P.x : -BIG);
        const double ly = wave_min(active ? P.y : BIG), hy = wave_max(active ? P.y : -BIG);
        const double lz = wave_min(active ? P.z : BIG), hz = wave_max(active ? P.z : -BIG);
        // the same on every lane: say so, so that it lives in scalar registers
        cx = uniform_d(0.5 * (lx + hx));
        cy = uniform_d(0.5 * (ly + hy));
        cz = uniform_d(0.5 * (lz + hz));
    }
    float qx = 0, qy = 0, qz = 0, hQ = 0;
    bool wave_unsafe = false;
    if constexpr (MIXED) {
        const double rx = wrapd(P.x - cx, Lx), ry = wrapd(P.y - cy, Ly), rz = wrapd(P.z - cz, Lz);
        qx = (float)rx;
        qy = (float)ry;
        qz = (float)rz;
        const double hql = active ? fmax(fabs(rx), fmax(fabs(ry), fabs(rz))) : 0.0;
        // the error bounds of this lane's keys need only its own offset from the origin
        hQ = (float)hql * (1.0f + 4 * U32);
        const double hq = wave_max(hql);
        // the FP32 frame is valid while search radius + query spread < L/2
        wave_unsafe = hq >= 0.25 * fmin(Lx, fmin(Ly, Lz));
    }
    for (int e = 0; e < KHCAP; ++e) HSET(e, e < KH ? (KeyT)INFINITY : (KeyT)-INFINITY, -1);
    KeyT tau = (active && !wave_unsafe) ? (KeyT)INFINITY : (KeyT)-1;
    // Morton-order seed window of this lane: [q-Wd, q+Wd]; the traversal skips it
    const int Wd = K / 2 + AHOP_SEED_PAD;
    int seed_lo = 1, seed_hi = 0;
    if (AHOP_WAVE_SEED && !qlist) {
        // every lane seeds from the whole window of the wave (wave-uniform; a lane without a
        // search accepts nothing, its tau being -1)
        seed_lo = max(0, wave * WAVE - Wd);
        seed_hi = min(A.n - 1, wave * WAVE + WAVE - 1 + Wd);
    } else if (active && !wave_unsafe) {
        seed_lo = max(0, q - Wd);
        seed_hi = min(A.n - 1, q + Wd);
    }
    // slots that may lie in some lane's seed window or be some lane's query (wave-uniform)
    const int wseed_lo = qlist ? 0 : max(0, wave * WAVE - Wd);
    const int wseed_hi = qlist ? A.n - 1 : min(A.n - 1, wave * WAVE + WAVE - 1 + Wd);
    const float pxf = (float)P.x, pyf = (float)P.y, pzf = (float)P.z;
    const float Lxf = (float)Lx, Lyf = (float)Ly, Lzf = (float)Lz;
    const float mLx = 4.0f * U32 * Lxf, mLy = 4.0f * U32 * Lyf, mLz = 4.0f * U32 * Lzf;
    const float Lmin = fminf(Lxf, fminf(Lyf, Lzf));
    const int qmid = qlist ? __builtin_amdgcn_readfirstlane(q) : wave * WAVE + WAVE / 2;
    uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;   // accepted-candidate masks of staged buckets
    int f0 = 0, f1 = 0, f2 = 0, f3 = 0;         // first slot of each staged bucket
    int nbuf = 0;
    // heap entries written so far by the seeding (the per-lane gather of a query list
    // inserts into the +inf-filled heap directly)
    int filled = (qlist || (!MIXED && !AHOP_FP64_FILL)) ? KH : 0;
    unsigned long long st_nodes = 0, st_buckets = 0, st_cand = 0, st_steps = 0;
    unsigned st_ins = 0;   // per lane

    // FP32 keys use fused multiply-adds: one rounding per step, so the error bound of
    // fp32_key_err (written for separate roundings) still holds
#define KNN_KEY(c, out)                                                                        \
    do {                                                                                       \
        if constexpr (MIXED) {                                                                 \
            const float dx_ = (c).x - qx, dy_ = (c).y - qy, dz_ = (c).z - qz;                  \
            (out) = __builtin_fmaf(dz_, dz_, __builtin_fmaf(dy_, dy_, dx_ * dx_));             \
        } else {                                                                               \
            (out) = d2_ref(P.x, P.y, P.z, (c).x, (c).y, (c).z, Lx, Ly, Lz);                    \
        }                                                                                      \
    } while (0)

#define KNN_INSERT(d, id)                                                                      \
    do {                                                                                       \
        HSIFT(KH, (d), (id), 0, SIFT_CHECKED);                                                 \
        ++st_ins;                                                                              \
        tau = HK(0);                                                                           \
    } while (0)

    // Take this lane's next pending candidate (first staged bucket with a set bit, lowest
    // bit first): its key and sorted slot.
#define KNN_PICK(dout, idout)                                                                  \
    do {                                                                                       \
        const bool s0 = m0 != 0;                                                               \
        const bool s1 = !s0 && (NBUF == 2 || m1 != 0);                                         \
        const bool s2 = NBUF > 2 && !s0 && !s1 && (NBUF == 3 || m2 != 0);                      \
        const bool s3 = NBUF > 3 && !s0 && !s1 && !s2;                                         \
        uint64_t mk = s0 ? m0 : (s1 ? m1 : (s2 ? m2 : m3));                                    \
        const int fsel = s0 ? f0 : (s1 ? f1 : (s2 ? f2 : f3));                                 \
        const int bsel = s0 ? 0 : (s1 ? 1 : (s2 ? 2 : 3));                                     \
        const int j_ = __ffsll((long long)mk) - 1;                                             \
        mk &= mk - 1;                                                                          \
        m0 = s0 ? mk : m0;                                                                     \
        m1 = s1 ? mk : m1;                                                                     \
        m2 = s2 ? mk : m2;                                                                     \
        m3 = s3 ? mk : m3;                                                                     \
        KNN_KEY(stage_at<MIXED>(sx, sy, sz, stage, bsel * WAVE + j_), dout);                   \
        (idout) = fsel + j_;                                                                   \
    } while (0)

    // Bottom-up heapify of this lane's KH entries (mixed: the register children that have
    // children, then the root).
#define KNN_HEAPIFY()                                                                          \
    do {                                                                                       \
        if constexpr (MIXED) {                                                                 \
            for (int e_ = RC; e_ >= 0; --e_)                                                   \
                if (e_ == 0 || rt_child0<RC>(e_) < KH) HSIFT(KH, HK(e_), HI(e_), e_, SIFT_CHECKED); \
        } else {                                                                               \
            for (int e_ = (KH - 2) / HEAP_D; e_ >= 0; --e_)                                    \
                HSIFT(KH, HK(e_), HI(e_), e_, SIFT_CHECKED);                                   \
        }                                                                                      \
    } while (0)

    // Insert every pending accepted candidate lane by lane: the wave pays for the busiest
    // lane's total over all staged buckets, not for each bucket's busiest lane.  While the
    // lane's heap is still being filled (seeding, FILL true), candidates are written straight
    // into the next free entry and the full array is heapified once, instead of sifting each
    // one; the traversal's flushes (FILL false) carry no fill code.
#define KNN_FLUSH(FILL)                                                                        \
    do {                                                                                       \
        while (__any((m0 | m1 | m2 | m3) != 0)) {                                             \
            ++st_steps;                                                                        \
            if ((m0 | m1 | m2 | m3) != 0) {                                                    \
                KeyT d_;                                                                       \
                int id_;                                                                       \
                KNN_PICK(d_, id_);                                                             \
                if ((FILL) && filled < KH) {                                                   \
                    HSET(filled, d_, id_);                                                     \
                    ++st_ins;                                                                  \
                    if (++filled == KH) {                                                      \
                        KNN_HEAPIFY();                                                         \
                        tau = HK(0);                                                        \
                    }                                                                          \
                } else if (d_ < tau) {                                                         \
                    KNN_INSERT(d_, id_);                                                       \
                }                                                                              \
            }                                                                                  \
        }                                                                                      \
        nbuf = 0;                                                                              \
    } while (0)

    // One pending candidate of this lane as KNN_PICK takes it, or +inf when the lane has none
    // left (the stage index then stays in range: bucket 1, slot 63).
#define KNN_PICK_V(dout, idout)                                                                \
    do {                                                                                       \
        const bool any_ = (m0 | m1) != 0;                                                      \
        const bool s0_ = m0 != 0;                                                              \
        uint64_t mk_ = s0_ ? m0 : m1;                                                          \
        const int fsel_ = s0_ ? f0 : f1;                                                       \
        const int bsel_ = s0_ ? 0 : 1;                                                         \
        const int j_ = __ffsll((long long)mk_) - 1;                                            \
        mk_ &= mk_ - 1;                                                                        \
        m0 = s0_ ? mk_ : m0;                                                                   \
        m1 = s0_ ? m1 : mk_;                                                                   \
        KNN_KEY(stage_at<MIXED>(sx, sy, sz, stage, bsel_ * WAVE + j_), dout);                  \
        (dout) = any_ ? (dout) : (KeyT)INFINITY;                                               \
        (idout) = fsel_ + j_;                                                                  \
    } while (0)

#define KNN_CSWAP(a, ia, b, ib)                                                                \
    do {                                                                                       \
        const bool sw_ = (b) < (a);                                                            \
        const KeyT ta_ = sw_ ? (b) : (a), tb_ = sw_ ? (a) : (b);                               \
        const int tia_ = sw_ ? (ib) : (ia), tib_ = sw_ ? (ia) : (ib);                          \
        (a) = ta_;                                                                             \
        (b) = tb_;                                                                             \
        (ia) = tia_;                                                                           \
        (ib) = tib_;                                                                           \
    } while (0)

    // Traversal flush of the mixed kernel: every step takes up to PICKW pending candidates of
    // each lane at once (their stage reads in flight together), sorts them, and inserts them
    // smallest first while they beat the lane's current k-th key.
#define KNN_FLUSH_W()                                                                          \
    do {                                                                                       \
        while (__any((m0 | m1) != 0)) {                                                        \
            ++st_steps;                                                                        \
            KeyT a0_, a1_, a2_ = (KeyT)INFINITY, a3_ = (KeyT)INFINITY;                         \
            int i0_, i1_, i2_ = 0, i3_ = 0;                                                    \
            KNN_PICK_V(a0_, i0_);                                                              \
            KNN_PICK_V(a1_, i1_);                                                              \
            if constexpr (AHOP_PICKW > 2) {                                                    \
                KNN_PICK_V(a2_, i2_);                                                          \
                KNN_PICK_V(a3_, i3_);                                                          \
                KNN_CSWAP(a0_, i0_, a1_, i1_);                                                 \
                KNN_CSWAP(a2_, i2_, a3_, i3_);                                                 \
                KNN_CSWAP(a0_, i0_, a2_, i2_);                                                 \
                KNN_CSWAP(a1_, i1_, a3_, i3_);                                                 \
                KNN_CSWAP(a1_, i1_, a2_, i2_);                                                 \
            } else {                                                                           \
                KNN_CSWAP(a0_, i0_, a1_, i1_);                                                 \
            }                                                                                  \
            while (__any(a0_ < tau)) {                                                         \
                if (a0_ < tau) KNN_INSERT(a0_, i0_);                                           \
                a0_ = a1_;                                                                     \
                i0_ = i1_;                                                                     \
                if constexpr (AHOP_PICKW > 2) {                                                \
                    a1_ = a2_;                                                                 \
                    i1_ = i2_;                                                                 \
                    a2_ = a3_;                                                                 \
                    i2_ = i3_;                                                                 \
                    a3_ = (KeyT)INFINITY;                                                      \
                } else {                                                                       \
                    a1_ = (KeyT)INFINITY;                                                      \
                }                                                                              \
            }                                                                                  \
        }                                                                                      \
        nbuf = 0;                                                                              \
    } while (0)
    // A seeding flush: once every lane's heap is full (the direct fill), two pending
    // candidates per lane per step as in the traversal.
#define KNN_FLUSH_S()                                                                          \
    do {                                                                                       \
        if (__all(filled >= KH)) KNN_FLUSH_T();                                                \
        else KNN_FLUSH(true);                                                                  \
    } while (0)
#define KNN_FLUSH_T()                                                                          \
    do {                                                                                       \
        if constexpr (MIXED && NBUF == 2 && AHOP_PICKW > 1) KNN_FLUSH_W();                     \
        else KNN_FLUSH(false);                                                                 \
    } while (0)

    // Evaluate the staged candidates j = 0..cnt-1 of buffer b (slots fst + j) and set bit j of
    // mask_ where ACCEPT holds (ACCEPT sees the key d_ and the slot sj_).  The mixed kernel
    // reads the x, y, z arrays four candidates at a time; the bits of the stale entries past
    // cnt (every buffer has WAVE entries) are cleared afterwards.
#define KNN_SCAN(b, fst, cnt, ACCEPT)                                                          \
    do {                                                                                       \
        if constexpr (MIXED) {                                                                 \
            /* the two 32-candidate halves build 32-bit masks (one select and one or a bit) */ \
            uint32_t hm_[2] = {0u, 0u};                                                        \
            _Pragma("unroll") for (int h_ = 0; h_ < 2; ++h_) {                                 \
                for (int j4_ = 32 * h_; j4_ < min((cnt), 32 * h_ + 32); j4_ += 4) {            \
                    const float4 X_ = *reinterpret_cast<const float4 *>(sx + (b) * WAVE + j4_); \
                    const float4 Y_ = *reinterpret_cast<const float4 *>(sy + (b) * WAVE + j4_); \
                    const float4 Z_ = *reinterpret_cast<const float4 *>(sz + (b) * WAVE + j4_); \
                    const float xs_[4] = {X_.x, X_.y, X_.z, X_.w};                             \
                    const float ys_[4] = {Y_.x, Y_.y, Y_.z, Y_.w};                             \
                    const float zs_[4] = {Z_.x, Z_.y, Z_.z, Z_.w};                             \
                    _Pragma("unroll") for (int t_ = 0; t_ < 4; ++t_) {                         \
                        const int j_ = j4_ + t_;                                               \
                        const int sj_ = (fst) + j_;                                            \
                        KeyT d_;                                                               \
                        KNN_KEY(make_float4(xs_[t_], ys_[t_], zs_[t_], 0.0f), d_);             \
                        if (ACCEPT) hm_[h_] |= 1u << (j_ - 32 * h_);                           \
                    }                                                                          \
                }                                                                              \
            }                                                                                  \
            mask_ = ((uint64_t)hm_[1] << 32) | hm_[0];                                         \
            if ((cnt) < WAVE) mask_ &= (1ull << (cnt)) - 1ull;                                 \
        } else {                                                                               \
            for (int j_ = 0; j_ < (cnt); ++j_) {                                               \
                const int sj_ = (fst) + j_;                                                    \
                KeyT d_;                                                                       \
                KNN_KEY(stage[(b) * WAVE + j_], d_);                                           \
                if (ACCEPT) mask_ |= (1ull << j_);                                             \
            }                                                                                  \
        }                                                                                      \
    } while (0)

    // Stage sorted slots [f, f+cnt) (cnt <= 64) and record each lane's accepted subset;
    // seeding accepts only the lane's own window, the traversal everything outside it.  The
    // seeding flushes a full stage before staging the next bucket, the traversal right after
    // filling it.
#define KNN_STAGE(fst, cnt, seeding)                                                           \
    do {                                                                                       \
        if ((seeding) && nbuf == NBUF) KNN_FLUSH_S();                                          \
        __syncthreads();                                                                       \
        if (lane < (cnt)) {                                                                    \
            const double4 c_ = spos[(fst) + lane];                                             \
            if constexpr (MIXED) {                                                             \
                sx[nbuf * WAVE + lane] = (float)wrapd(c_.x - cx, Lx);                          \
                sy[nbuf * WAVE + lane] = (float)wrapd(c_.y - cy, Ly);                          \
                sz[nbuf * WAVE + lane] = (float)wrapd(c_.z - cz, Lz);                          \
            } else {                                                                           \
                stage[nbuf * WAVE + lane] = c_;                                                \
            }                                                                                  \
        }                                                                                      \
        __syncthreads();                                                                       \
        uint64_t mask_ = 0;                                                                    \
        if ((seeding) || ((fst) <= wseed_hi && (fst) + (cnt) - 1 >= wseed_lo)) {               \
            KNN_SCAN(nbuf, fst, cnt,                                                           \
                     d_ < tau && sj_ != q &&                                                   \
                         ((seeding) ? (sj_ >= seed_lo && sj_ <= seed_hi)                       \
                                    : !(sj_ >= seed_lo && sj_ <= seed_hi)));                   \
        } else {                                                                               \
            /* the bucket lies outside every lane's seed window (and holds no query) */       \
            KNN_SCAN(nbuf, fst, cnt, d_ < tau);                                                \
        }                                                                                      \
        m0 = nbuf == 0 ? mask_ : m0;                                                           \
        m1 = nbuf == 1 ? mask_ : m1;                                                           \
        if constexpr (NBUF > 2) m2 = nbuf == 2 ? mask_ : m2;                                   \
        if constexpr (NBUF > 3) m3 = nbuf == 3 ? mask_ : m3;                                   \
        f0 = nbuf == 0 ? (fst) : f0;                                                           \
        f1 = nbuf == 1 ? (fst) : f1;                                                           \
        if constexpr (NBUF > 2) f2 = nbuf == 2 ? (fst) : f2;                                   \
        if constexpr (NBUF > 3) f3 = nbuf == 3 ? (fst) : f3;                                   \
        ++nbuf;                                                                                \
    } while (0)

    // visit test of node c for this lane (out), with the node's box distance (bdout)
#define KNN_NEED(c, out, bdout)                                                                \
    do {                                                                                       \
        const float bd_ = box_dist2_lower((c), spos, nfbox, pxf, pyf, pzf, Lxf, Lyf, Lzf,     \
                                          mLx, mLy, mLz);                                      \
        (bdout) = bd_;                                                                         \
        if constexpr (MIXED) {                                                                 \
            (out) = active && bd_ < tn;                                                        \
        } else {                                                                               \
            (out) = active && (double)bd_ < tau;                                               \
        }                                                                                      \
    } while (0)

    // visit test of a child whose box comes from its parent's record (a: cx, cy, cz, hx;
    // b: hy, hz, ...)
#define KNN_NEED_AT(a, b, out, bdout)                                                          \
    do {                                                                                       \
        const float bd_ = box_dist2_at((a).x, (a).y, (a).z, (a).w, (b).x, (b).y, pxf, pyf,      \
                                       pzf, Lxf, Lyf, Lzf, mLx, mLy, mLz);                      \
        (bdout) = bd_;                                                                         \
        if constexpr (MIXED) {                                                                 \
            (out) = active && bd_ < tn;                                                        \
        } else {                                                                               \
            (out) = active && (double)bd_ < tau;                                               \
        }                                                                                      \
    } while (0)

    // ---- seed: the wave's seed windows cover the contiguous slots [qb-Wd, qb+63+Wd]
    if (!qlist) {
        const int qb = wave * WAVE;
        const int s0 = max(0, qb - Wd), s1 = min(A.n - 1, qb + WAVE - 1 + Wd);
        for (int f = s0; f <= s1; f += WAVE) {
            const int cnt = min(WAVE, s1 - f + 1);
            KNN_STAGE(f, cnt, true);
            if constexpr (MIXED) {
                // Two full buckets of the window are staged, and every lane with a
                // search has accepted all of them but its own slot: KH of them go straight
                // into its heap, its own slot (if among them) with key +inf, which the first
                // insertion evicts, and the fill's picks of them are dropped.  A lane takes
                // the KH slots centred on its own as far as the two buckets allow (the first
                // KH slots for every lane: 179 insertions per lane at the 16.8M box, 96.98
                // ms; centred: 169, 96.64 ms).
                if (nbuf == 2 && f == s0 + WAVE && cnt == WAVE && KH <= 2 * WAVE) {
                    const int off = q - s0;
                    const uint64_t self0 = off < WAVE ? 1ull << off : 0ull;
                    const uint64_t self1 = (off >= WAVE && off < 2 * WAVE) ? 1ull << (off - WAVE) : 0ull;
                    if (m0 == ~self0 && m1 == ~self1) {
                        // each lane takes the KH staged slots centred on its own (as near as
                        // the two buckets allow): relative slots [rs, rs + KH)
                        const int rs = min(max(off - KH / 2, 0), 2 * WAVE - KH);
#pragma unroll
                        for (int e = 0; e < RT; ++e) {
                            KeyT d;
                            const int r = rs + e;
                            KNN_KEY(make_float4(sx[r], sy[r], sz[r], 0.0f), d);
                            // (the host launches capacities of 32 or more only for k > 16,
                            // so there KH > RT and the size test can go)
                            if (KCAP >= 32 || e < KH) HSET(e, r == off ? (KeyT)INFINITY : d, s0 + r);
                        }
                        for (int e = RT; e < KH; ++e) {
                            KeyT d;
                            const int r = rs + e;
                            KNN_KEY(make_float4(sx[r], sy[r], sz[r], 0.0f), d);
                            d = r == off ? (KeyT)INFINITY : d;
                            hp[(e - RT) * WAVE + lane] = make_uint2(__float_as_uint(d), (uint32_t)(s0 + r));
                        }
                        filled = KH;
                        st_ins += KH;
                        KNN_HEAPIFY();
                        tau = HK(0);
                        m0 &= ~bits_in(rs, rs + KH);
                        m1 &= ~bits_in(rs - WAVE, rs + KH - WAVE);
                    }
                }
            }
        }
    } else {
        for (int j = seed_lo; j <= seed_hi; ++j) {
            if (j == q) continue;
            const double4 c = spos[j];
            KeyT d;
            if constexpr (MIXED)
                KNN_KEY(make_float4((float)wrapd(c.x - cx, Lx), (float)wrapd(c.y - cy, Ly),
                                    (float)wrapd(c.z - cz, Lz), 0.0f),
                        d);
            else
                KNN_KEY(c, d);
            if (d < tau) KNN_INSERT(d, j);
        }
    }
    KNN_FLUSH_S();
    if (filled < KH) {
        // a clipped window (array ends) or an idle lane: the unfilled entries are +inf
        KNN_HEAPIFY();
        filled = KH;
        tau = (active && !wave_unsafe) ? HK(0) : (KeyT)-1;
    }

    // ---- wave-coherent traversal of the binary radix tree, nearest child first
    int sp = 1;
    stack[0] = 0;
    bool stack_overflow = false;
    const bool root_is_bucket = (A.n <= A.bucket);
    // the node on top of the stack was pushed by the node just handled, and its visit test
    // from then (top_need) still holds: tau changes only when a bucket is staged
    bool top_fresh = false, top_need = false;
    while (sp > 0) {
        // the traversal is wave-uniform: say so, so that node data come through scalar loads
        const int node = __builtin_amdgcn_readfirstlane(stack[--sp]);
        // everything the visit reads, in one round of scalar loads: the node's links and
        // slot range, and its children's boxes and slot ranges
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
        float tn = 0.0f;
        if constexpr (MIXED) tn = tau + 2 * fp32_key_err(tau, hQ, Lmin);
        bool need = top_need;
        float bdn;
        if (!top_fresh) KNN_NEED(node, need, bdn);
        top_fresh = false;
        if (!__any(need)) continue;
        if (l - f + 1 <= A.bucket || (root_is_bucket && node == 0)) {
            // a bucket inside the wave-wide seed window was seen by every lane already
            if (AHOP_WAVE_SEED && !qlist && f >= wseed_lo && l <= wseed_hi) continue;
            const int cnt = l - f + 1;
            ++st_buckets;
            st_cand += cnt;
            KNN_STAGE(f, cnt, false);
            // flushed as soon as the stage is full, so that the tests of the nodes in
            // between see the new roots
            if (nbuf == NBUF) KNN_FLUSH_T();
            continue;
        }
        const int c1 = rec.x;
        const int c2 = rec.y;
        bool n1, n2;
        float bd1, bd2;
        KNN_NEED_AT(cb0, cb1, n1, bd1);
        KNN_NEED_AT(cb2, cb3, n2, bd2);
        top_fresh = true;
        const bool a1 = __any(n1), a2 = __any(n2);
        if (sp + 2 > STACK) {
            stack_overflow = true;   // reported after the loop (no global atomics inside it)
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
                stack[sp++] = c2;
                stack[sp++] = c1;
                top_need = n1;
            } else {
                stack[sp++] = c1;
                stack[sp++] = c2;
                top_need = n2;
            }
        } else if (a1) {
            stack[sp++] = c1;
            top_need = n1;
        } else if (a2) {
            stack[sp++] = c2;
            top_need = n2;
        } else {
            top_fresh = false;
        }
    }
    KNN_FLUSH_T();
    if (stack_overflow && lane == 0) atomicAdd(A.overflow, 1);
#undef KNN_NEED
#undef KNN_NEED_AT
#undef KNN_STAGE
#undef KNN_SCAN
#undef KNN_FLUSH
#undef KNN_FLUSH_W
#undef KNN_FLUSH_T
#undef KNN_FLUSH_S
#undef KNN_PICK_V
#undef KNN_CSWAP
#undef KNN_INSERT
#undef KNN_KEY

    if (A.stats) {
        unsigned long long ins = st_ins;
        for (int o = 32; o > 0; o >>= 1) ins += __shfl_xor(ins, o);
        if (lane == 0) {
            atomicAdd(&A.stats[KNN_STAT_NODES], st_nodes);
            atomicAdd(&A.stats[KNN_STAT_BUCKETS], st_buckets);
            atomicAdd(&A.stats[KNN_STAT_CANDIDATES], st_cand);
            atomicAdd(&A.stats[KNN_STAT_INSERT_STEPS], st_steps);
            atomicAdd(&A.stats[KNN_STAT_INSERTS], ins);
            atomicAdd(&A.stats[KNN_STAT_WAVES], 1ull);
        }
    }

    // ---- heap -> ascending order (in-place heapsort, same trip count on every lane)
    for (int end = KH - 1; end >= 1; --end) {
        const KeyT top = HK(0);
        const int32_t tid = HI(0);
        const KeyT mv = HK(end);
        const int32_t mid = HI(end);
        HSET(end, top, tid);
        HSIFT(end, mv, mid, 0, true);
    }
    if (!active) return;

    bool ok = !wave_unsafe;
    // entries that take part in the FP64 selection of the k nearest: all but the last heap
    // entry (the mixed heap holds k+KX), and no more than the other particles there are
    const int nsel = MIXED ? min(KH - 1, A.n - 1) : K;
    if constexpr (MIXED) {
        // certify that FP32 ranking cannot have left one of the k nearest out of the first
        // k+KX-1 entries: the last FP32 key must clear the k-th by twice the error bound E
        // of a key of that size (every particle of the true k nearest has a key below the
        // k-th key + 2E), and every particle within that radius must lie in the periodic
        // image the wave-centred frame picks (radius + query spread < L/2; a particle the
        // frame folds onto another image gets a key no smaller than its true distance, so
        // only the particles inside the radius need the right image).  An unfilled heap
        // (last key infinite) holds every other particle.
        const float lkx = HK(KH - 1), lk1 = HK(K - 1);
        const float E = fp32_key_err(lkx, hQ, Lmin);
        if (!(lkx == INFINITY) &&
            !(lkx - lk1 > 2 * E && E < 0.01f * lkx && sqrtf(lkx * 1.0001f) + hQ < 0.4999f * Lmin))
            ok = false;
    }
    if (!ok) {
        A.fb_list[atomicAdd(A.fb_count, 1)] = q;
        return;
    }
    // ---- density in FP64, reference order (omp_compute_density f90:2816-2838); the mixed
    // path recomputes every distance in the reference's FP64 arithmetic; where FP32 near-ties
    // left entries out of FP64 order (within the k, or an extra entry nearer than the k-th),
    // the lane re-sorts its nsel entries by the FP64 distances, keeps the first k and sums
    // again, so its densities are those of the FP64 path
    // the query's coordinates and mass, read again here rather than held in registers
    // through the search (the FP64 kernel's keys use them throughout)
    const double4 Q = MIXED ? spos[q] : P;
    double rho = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        double d2k;
        if constexpr (MIXED) {
            const double4 c = spos[HI(K - 1)];
            d2k = d2_ref(Q.x, Q.y, Q.z, c.x, c.y, c.z, Lx, Ly, Lz);
        } else {
            d2k = HK(K - 1);
        }
        const double r = sqrt(d2k) * 0.5;
        const double r3 = r * r * r;
        const double unsr = 1.0 / r;
        double contrib = 0.0;
        double prev = -1.0;
        bool sorted = true;
        for (int i = 0; i < K - 1; ++i) {
            const double4 c = spos[HI(i)];
            double d;
            if constexpr (MIXED) d = d2_ref(Q.x, Q.y, Q.z, c.x, c.y, c.z, Lx, Ly, Lz);
            else d = HK(i);
            sorted = sorted && (d >= prev);
            prev = d;
            contrib = contrib + c.w * spline_ref(sqrt(d) * unsr);
        }
        sorted = sorted && (d2k >= prev);
        if constexpr (MIXED) {
            for (int i = K; i < nsel; ++i) {
                const double4 c = spos[HI(i)];
                sorted = sorted && (d2_ref(Q.x, Q.y, Q.z, c.x, c.y, c.z, Lx, Ly, Lz) >= d2k);
            }
        }
        rho = A.xyzl * (contrib + Q.w) / (PI_REF * r3);
        if (!MIXED || sorted) break;
        // insertion sort of this lane's nsel entries by FP64 distance (stable on ties)
        for (int i = 1; i < nsel; ++i) {
            const int32_t id = HI(i);
            const KeyT ki = HK(i);
            const double4 ci = spos[id];
            const double di = d2_ref(Q.x, Q.y, Q.z, ci.x, ci.y, ci.z, Lx, Ly, Lz);
            int j = i - 1;
            while (j >= 0) {
                const double4 cj = spos[HI(j)];
                if (d2_ref(Q.x, Q.y, Q.z, cj.x, cj.y, cj.z, Lx, Ly, Lz) <= di) break;
                HSET(j + 1, HK(j), HI(j));
                --j;
            }
            HSET(j + 1, ki, id);
        }
    }
    A.rho_s[q] = rho;
    for (int t = 0; t < A.nhop; ++t) A.nbr[(int64_t)q * A.nhop + t] = HI(t);
}
#undef KNN_PICK
#undef KNN_HEAPIFY
#undef HK
#undef HI
#undef HSET
#undef HSIFT
