// Short-range Yukawa-Ewald PP forces and energy on a sub-cell list.
//
// Reference: particle_particle, yukawa_pp.py:6-139 (C twin yukawa_p3m.c:182-416).
// Same pair set (every pair with minimum-image distance r < rc, rc <= L/3),
// same pair terms (yukawa_pp.py:115-124), same displacement convention
// d = x_i - (x_j + shift) (:105-107), same cutoff r < rc (:109-112, tested as
// d.d < rc^2; the two differ only for r within an ulp of rc).
// Only the organisation of the work and the summation order differ:
//
// * Pair terms.  With b = kappa / (2 alpha) the reference's energy and force
//   terms are phi(r) = g1(r) / r and f(r) = g2(r) / r^2 with
//     g1 = 1/2 [e^{kr} erfc(ar+b) + e^{-kr} erfc(ar-b)],
//     g2 = 1/2 [e^{kr} erfc(ar+b)(1-kr) + e^{-kr} erfc(ar-b)(1+kr)]
//          + (2a/sqrt(pi)) r e^{-(ar)^2 - b^2}
//   (e^{-(ar+b)^2} e^{kr} = e^{-(ar-b)^2} e^{-kr} = e^{-(ar)^2-b^2}).  g1 and g2
//   are entire functions of r, so on [0, rc] they are represented by
//   piecewise Chebyshev series (equal pieces, degree kChebDeg) fitted per
//   context on the host from long-double evaluations and checked there to
//   <= 4e-15 of max|g| (pair_table.hip).  A pair then costs one reciprocal
//   square root and two Clenshaw recurrences instead of a square root, a
//   division, three exponentials and two complementary error functions.
// * Cells.  One workgroup per reference cell = a 2x2x2 block of half-width
//   sub-cells.  The 6x6x6 sub-cells around the block are staged in LDS with
//   their periodic image shift applied, as 36 rows of 6 sub-cells (one or two
//   contiguous runs of the sorted particle array each).
// * Sweep.  A wave handles two i at a time, one per 32-lane half.  Each half
//   sweeps the 5 planes of its i's 5x5x5 sub-cell stencil, each plane one
//   contiguous LDS range (its 5 rows are consecutive tile rows; tabulated
//   once per home sub-cell), testing two batches of 32 candidates per step;
//   in-cutoff j are compacted into the half's LDS queue with a ballot and mbcnt, and the
//   pair terms run only on queue batches, so lanes stay busy despite the
//   ~25% in-cutoff rate of a cell stencil.
// * Each pair is evaluated from both ends (no Newton-3 scatter, no atomics);
//   i's force and energy are summed across its half-wave (DPP moves in
//   FP32, shuffles in FP64).
//   The energy is halved at the end: U_s_r = 1/2 sum_i sum_{j != i} phi.
// * A tile larger than the LDS capacity is processed in several passes over
//   its rows (i's force accumulates in global memory across passes).
// * Mixed precision (k_pp<float>): each tile coordinate, taken relative to
//   the centre of the home block in FP64, is stored as an FP32 pair hi + lo
//   (hi = fl32(c), lo = fl32(c - hi)).  The candidate sweep reads only hi;
//   a queued pair's displacement is (xi_hi - xj_hi) + (xi_lo - xj_lo): the
//   hi difference of two nearby coordinates is exact (Sterbenz), so the
//   displacement keeps FP32 relative accuracy even for pairs far closer than
//   the hi rounding (2^-24 of 1.5 cells).  The pair terms use FP32 tables
//   (degree kChebDegF, fitted to 1e-7 of max|g|, evaluated in the power
//   basis by Horner) and FP32 arithmetic; i's FP32 half-wave sums are
//   widened to FP64 for its acc write and the block energy.  The self pair
//   (hi and lo identical: zero displacement) and coincident particles are
//   dropped by the pair test.
#include "mdf_internal.h"

namespace mdf {

template <class T>
struct Vec2;
template <>
struct Vec2<double> {
    using t = double2;
};
template <>
struct Vec2<float> {
    using t = float2;
};

template <class T>
struct PPArgs {
    using T2 = typename Vec2<T>::t;
    Geom g;
    const double4 *__restrict__ spos;
    const int *__restrict__ cell_start;
    const T2 *__restrict__ cheb;       // [piece][k] = (g1 coeff, g2 coeff)
    double *__restrict__ acc;          // sorted (n, 3)
    double *__restrict__ part;         // per-block energy partial
    int *__restrict__ stats;
    int tile_max;
    int npiece;                        // pieces of the r-indexed table
    int ntab;                          // pieces of the table staged in LDS (cheb)
    T piece_w, inv_piece_w;
    // FP32: tile coordinates are in units of 1 / sigma (sigma = 2^(E/2) / rc,
    // close to 1), so a pair's s' = sigma^2 r^2 and the cutoff is s' = s_top =
    // 2^E exactly; cheb is the s-indexed table (SPieces) in s'; the r-indexed
    // table, for pairs closer than s_lo, is read from global memory
    const T2 *__restrict__ rtab;
    unsigned s_base;
    int s_shift;
    float s_lo, s_top, s_tc;
    double qscale;                     // sigma
    float qscale_f, inv_q;             // sigma and 1 / sigma in FP32
    T rc2;                             // rc^2, the pair cutoff (FP32: s_top)
    T rc2m;                            // candidate test radius^2 (rc slightly widened)
    // Verlet pair lists (null vl_flags: off).  vl_flags[3] is the step's mode:
    // 1 = build (the sweep tests candidates within the list radius, evaluates
    // those within rc and writes every candidate it queued to the list of its
    // i), 2 = evaluate the lists.  A build that cannot list a block (a tile in
    // several passes, a list over vl_cap) sets vl_flags[1], which makes the
    // next step build again.  On a list step spos and cell_start are the
    // frozen order the lists index (the binning keeps it until the next build).
    int *vl_flags;
    unsigned short *__restrict__ vl_list;  // vl_cap entries per sorted slot: tile slots of its j
    int *__restrict__ vl_cnt;               // entries per sorted slot
    int vl_cap;
    T rcl2;                                 // list radius^2: (rc + skin)^2, slightly widened (FP32: times sigma^2)
};

template <class T>
struct ChebDeg;
template <>
struct ChebDeg<double> {
    static constexpr int v = kChebDeg;
};
template <>
struct ChebDeg<float> {
    static constexpr int v = kChebDegF;
};

constexpr int kGroups = 2;             // i per wave
constexpr int kGL = 64 / kGroups;      // lanes per i
constexpr int kQCap = 2 * kGL;         // queue entries per group

// Clenshaw sums of both series of one piece at t in [-1, 1].
template <class T, class C>
__device__ __forceinline__ void cheb2(C c, T t, T &g1, T &g2) {
    constexpr int D = ChebDeg<T>::v;
    const T t2 = t + t;
    T a1 = 0, a2 = 0, b1 = 0, b2 = 0;
#pragma unroll
    for (int k = D; k >= 1; --k) {
        const T cx = c[k].x, cy = c[k].y;
        const T na = fma(t2, a1, cx - a2);
        const T nb = fma(t2, b1, cy - b2);
        a2 = a1;
        a1 = na;
        b2 = b1;
        b1 = nb;
    }
    g1 = fma(t, a1, c[0].x - a2);
    g2 = fma(t, b1, c[0].y - b2);
}

// Horner sums of both power series of one piece at t in [-1, 1] (FP32 tables).
template <class T, int D = ChebDeg<T>::v, class C>
__device__ __forceinline__ void horner2(C c, T t, T &g1, T &g2) {
    T a = c[D].x, b = c[D].y;
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
        a = fma(a, t, (T)c[k].x);
        b = fma(b, t, (T)c[k].y);
    }
    g1 = a;
    g2 = b;
}

template <class T>
__device__ __forceinline__ T rsqrt_t(T x);
template <>
__device__ __forceinline__ double rsqrt_t<double>(double x) {
    return rsqrt(x);
}
template <>
__device__ __forceinline__ float rsqrt_t<float>(float x) {
    // v_rsq_f32 (1 ulp) without rsqrtf's denormal rescaling: a pair's r^2 is a
    // normal number (positions are O(L), rc^2 >> FLT_MIN)
    return __builtin_amdgcn_rsqf(x);
}

template <class T>
__device__ __forceinline__ void pp_pair_cheb(const PPArgs<T> &P, const typename Vec2<T>::t *__restrict__ tab, T dx,
                                             T dy, T dz, T &fx, T &fy, T &fz, T &u) {
    constexpr int D = ChebDeg<T>::v;
    const T r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < P.rc2)) return;  // r < rc (yukawa_pp.py:112) tested as r^2 < rc^2
    if (sizeof(T) == 4 && !(r2 > (T)0)) return;
    const T ir = rsqrt_t<T>(r2);
    const T r = r2 * ir;
    // r > 0, so p >= 0: only the top needs a clamp (r within rounding of rc)
    const int p = min((int)(r * P.inv_piece_w), P.npiece - 1);
    const T t = (r - p * P.piece_w) * ((T)2 * P.inv_piece_w) - (T)1;
    // the table through an LDS-typed pointer: 32-bit address arithmetic
    using LdsT2 = __attribute__((address_space(3))) const typename Vec2<T>::t;
    LdsT2 *c = (LdsT2 *)tab + __umul24((unsigned)p, D + 1);  // p < 2^24: a full-rate 24-bit multiply
    T g1, g2;
    if constexpr (sizeof(T) == 4) horner2<T>(c, t, g1, g2);  // power basis (FP32 table)
    else cheb2<T>(c, t, g1, g2);
    u += g1 * ir;
    const T fr = g2 * ir * ir * ir;
    fx += fr * dx;
    fy += fr * dy;
    fz += fr * dz;
}

// The same pair terms without branches: a pair that is not ok, outside the
// cutoff or at r = 0 is evaluated at r = 1 and contributes nothing (selects,
// not multiplications, so no inf or NaN leaks in).  Two of these in a row
// have no control dependence, so their long latency chains interleave.  The
// arithmetic of a contributing pair is pp_pair_cheb's, operation for
// operation.
template <class T>
__device__ __forceinline__ void pp_pair_masked(const PPArgs<T> &P, const typename Vec2<T>::t *__restrict__ tab,
                                               bool ok, T dx, T dy, T dz, T &fx, T &fy, T &fz, T &u) {
    constexpr int D = ChebDeg<T>::v;
    const T r2 = dx * dx + dy * dy + dz * dz;
    const bool in = ok && r2 < P.rc2 && r2 > (T)0;
    const T r2s = in ? r2 : (T)1;
    const T ir = rsqrt_t<T>(r2s);
    const T r = r2s * ir;
    const int p = min((int)(r * P.inv_piece_w), P.npiece - 1);
    const T t = (r - p * P.piece_w) * ((T)2 * P.inv_piece_w) - (T)1;
    using LdsT2 = __attribute__((address_space(3))) const typename Vec2<T>::t;
    LdsT2 *c = (LdsT2 *)tab + __umul24((unsigned)p, D + 1);
    T g1, g2;
    if constexpr (sizeof(T) == 4) horner2<T>(c, t, g1, g2);
    else cheb2<T>(c, t, g1, g2);
    const T uu = in ? g1 * ir : (T)0;
    const T fr = in ? g2 * ir * ir * ir : (T)0;
    u += uu;
    fx = fma(fr, dx, fx);
    fy = fma(fr, dy, fy);
    fz = fma(fr, dz, fz);
}

// FP32 pair terms from the s-indexed table (pair_table.hip).  The
// displacement (dx, dy, dz) is in units of 1 / sigma, s' = |d|^2, and the
// table holds phi and F / sigma as functions of s', so F dx is the force in
// length units.  The piece and its local variable come from the float bits
// of s' clamped into [s_lo, s_top] by one med3: s' at or beyond the cutoff
// (s_top = sigma^2 rc^2 = 2^E) reads the all-zero last piece, which is the
// cutoff test (candidates beyond rc of a build queue or a list, the far
// dummy slot, pairs that are not ok evaluated at a far displacement).  Pairs
// with s' < s_lo (r < 1/2 .. 1/sqrt(2): the release of a random start, never
// the liquid) and coincident ions take the r-indexed table from global memory
// behind the one branch.  The sweep's queue holds i itself (s' = 0): that
// branch drops it too.
// s_top_v: P.s_top held in a VGPR by the caller (a med3 takes one SGPR operand).
__device__ __forceinline__ void pp_pair_f(const PPArgs<float> &P, unsigned tbase, float s_top_v, float dx, float dy,
                                          float dz, float &fx, float &fy, float &fz, float &u) {
    constexpr int D = kChebDegF;  // the r-indexed table of the close-pair path
    constexpr int DS = kSDegF;    // the s-indexed table in LDS
    const float s2 = dx * dx + dy * dy + dz * dz;
    const unsigned b = __float_as_uint(__builtin_amdgcn_fmed3f(s2, P.s_lo, s_top_v));
    const float t = __uint_as_float((b & ((1u << P.s_shift) - 1u)) | 0x3f800000u) - P.s_tc;
    // piece address: tbase is the table's LDS address less base pieces
    using LdsF2 = __attribute__((address_space(3))) const float2;
    LdsF2 *c = (LdsF2 *)(size_t)(__umul24(b >> P.s_shift, kSStride * (unsigned)sizeof(float2)) + tbase);
    float phi, F;
    horner2<float, DS>(c, t, phi, F);
    if (s2 < P.s_lo) {
        const float ir = rsqrt_t<float>(s2);  // 1 / (sigma r)
        const float r = s2 * ir * P.inv_q;
        const int q = min((int)(r * P.inv_piece_w), P.npiece - 1);
        const float tq = (r - q * P.piece_w) * (2.0f * P.inv_piece_w) - 1.0f;
        float g1, g2;
        horner2<float>(P.rtab + q * (D + 1), tq, g1, g2);
        const float irr = ir * P.qscale_f;  // 1 / r
        phi = g1 * irr;
        F = g2 * irr * irr * irr * P.inv_q;
        if (!(s2 > 0.0f)) phi = F = 0.0f;  // i itself, coincident ions: no pair term
    }
    u += phi;
    fx = fmaf(F, dx, fx);
    fy = fmaf(F, dy, fy);
    fz = fmaf(F, dz, fz);
}

// The table part of pp_pair_f alone (no branch); s' returned for the caller's
// close-pair bookkeeping.
__device__ __forceinline__ void pp_pair_tab(const PPArgs<float> &P, unsigned tbase, float s_top_v, float dx, float dy,
                                            float dz, float &fx, float &fy, float &fz, float &u, float &s2) {
    s2 = dx * dx + dy * dy + dz * dz;
    const unsigned b = __float_as_uint(__builtin_amdgcn_fmed3f(s2, P.s_lo, s_top_v));
    const float t = __uint_as_float((b & ((1u << P.s_shift) - 1u)) | 0x3f800000u) - P.s_tc;
    using LdsF2 = __attribute__((address_space(3))) const float2;
    LdsF2 *c = (LdsF2 *)(size_t)(__umul24(b >> P.s_shift, kSStride * (unsigned)sizeof(float2)) + tbase);
    float phi, F;
    horner2<float, kSDegF>(c, t, phi, F);
    u += phi;
    fx = fmaf(F, dx, fx);
    fy = fmaf(F, dy, fy);
    fz = fmaf(F, dz, fz);
}
// Correction of a close pair (s' < s_lo) that pp_pair_tab evaluated at the
// clamped s' = s_lo: its r-indexed terms less the table's.
__device__ __forceinline__ void pp_pair_fix(const PPArgs<float> &P, unsigned tbase, float s_top_v, float dx, float dy,
                                            float dz, float &fx, float &fy, float &fz, float &u) {
    constexpr int D = kChebDegF;
    const float s2 = dx * dx + dy * dy + dz * dz;
    if (!(s2 < P.s_lo)) return;
    const unsigned b = __float_as_uint(__builtin_amdgcn_fmed3f(s2, P.s_lo, s_top_v));
    const float t = __uint_as_float((b & ((1u << P.s_shift) - 1u)) | 0x3f800000u) - P.s_tc;
    using LdsF2 = __attribute__((address_space(3))) const float2;
    LdsF2 *c = (LdsF2 *)(size_t)(__umul24(b >> P.s_shift, kSStride * (unsigned)sizeof(float2)) + tbase);
    float phi_t, F_t;
    horner2<float, kSDegF>(c, t, phi_t, F_t);
    const float ir = rsqrt_t<float>(s2);
    const float r = s2 * ir * P.inv_q;
    const int q = min((int)(r * P.inv_piece_w), P.npiece - 1);
    const float tq = (r - q * P.piece_w) * (2.0f * P.inv_piece_w) - 1.0f;
    float g1, g2;
    horner2<float>(P.rtab + q * (D + 1), tq, g1, g2);
    const float irr = ir * P.qscale_f;
    float phi = g1 * irr, F = g2 * irr * irr * irr * P.inv_q;
    if (!(s2 > 0.0f)) phi = F = 0.0f;
    u += phi - phi_t;
    fx = fmaf(F - F_t, dx, fx);
    fy = fmaf(F - F_t, dy, fy);
    fz = fmaf(F - F_t, dz, fz);
}

__device__ __forceinline__ void wrap_cell(int c, int n, double L, int &cw, double &shift) {
    if (c < 0) {
        cw = c + n;
        shift = -L;
    } else if (c >= n) {
        cw = c - n;
        shift = L;
    } else {
        cw = c;
        shift = 0.0;
    }
}

// image shift of tile sub-cell index c (one period at most) along an axis of n sub-cells
__device__ __forceinline__ double image_shift(int c, int n, double L) { return c < 0 ? -L : (c >= n ? L : 0.0); }

__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Sum of v over each 32-lane half of the wave, returned in every lane of the
// half.  DPP moves (VALU, no LDS): xor 1 and xor 2 by quad permutes, then the
// half-row and row mirrors (sums of 8, then 16 lanes), then row_bcast:15
// adds row 0 into row 1 and row 2 into row 3; one shuffle takes each half's
// total from its last lane.  All lanes must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
__device__ __forceinline__ float half_wave_sum(float v) {
    v = dpp_add<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_add<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_add<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_add<0x140, 0xF>(v);  // row_mirror
    v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    return __shfl(v, kGL - 1, kGL);
}

// Slots a list of cnt entries occupies with its padding: whole chunks of
// 2 kGL entries and at least two padding slots (one list word) at the end.
// List layout: chunks of 2 kGL entries; dword l of a chunk holds its entries
// l (low half) and kGL + l (high half), so the first kGL entries of a chunk
// are the low halves of all its dwords.  Entry e sits at 16-bit slot
// list_pos(e); a list is padded with the dummy slot to a whole chunk.
__device__ __forceinline__ int list_pos(int e) {
    return (e & ~(2 * kGL - 1)) + 2 * (e & (kGL - 1)) + ((e / kGL) & 1);
}
__device__ __forceinline__ int list_padded(int cnt) { return (cnt + 2 * kGL - 1) & ~(2 * kGL - 1); }

// NW waves per workgroup share the tile: 4 for tiles up to 64 KB, 16 for
// larger ones (large rc), where one workgroup per CU would otherwise leave a
// single wave per SIMD.
template <class R, int NW, bool LISTS>
__device__ __forceinline__ void pp_body(const PPArgs<R> &P, const Blk &blk, unsigned char *ldsraw) {
    using R2 = typename Vec2<R>::t;
    constexpr bool kRel = sizeof(R) == 4;  // tile relative to the home block centre
    // degree of the table staged in LDS: FP32 the s-indexed pieces, FP64 the
    // r-indexed Chebyshev pieces
    constexpr int D = kRel ? kSDegF : ChebDeg<R>::v;
    // FP32 tile: hi and lo as two float4 (32 B a particle), or packed into a
    // float4 (hi x, y, z, lo x) and a float2 (lo y, z) (24 B) for the 12-wave
    // kernel, where the smaller tile is what lets two workgroups share a CU;
    // at equal occupancy the packed pair read measured ~10% slower
    constexpr bool kPack = kRel && NW == 12;
    const int T = P.tile_max;
    // FP64: three coordinate arrays (24 B a particle).  FP32: hi and lo as
    // two float4 per particle (32 B), one 16-byte LDS read each
    R *tx = (R *)ldsraw;
    R *ty = tx + T;
    R *tz = ty + T;
    float4 *th4 = (float4 *)ldsraw;                        // mixed: hi (x, y, z), packed: + lo x
    float4 *tl4 = th4 + T;                                 // mixed, unpacked: lo (x, y, z)
    float2 *tl2 = (float2 *)(th4 + T);                     // mixed, packed: lo (y, z)
    R2 *tab = kRel ? (kPack ? (R2 *)(tl2 + T) : (R2 *)(tl4 + T)) : (R2 *)(tz + T);  // npiece * (D + 1)
    constexpr int TS = kRel ? kSStride : D + 1;  // LDS stride of a table piece
    int *cum = (int *)(tab + P.ntab * TS);                 // 217 prefix offsets of the tile sub-cells
    int *queue = cum + 224;                                // NW * kGroups * kQCap
    __shared__ int tstart[216];  // sorted-array start of each tile sub-cell
    __shared__ int2 rng[8][5];   // per home sub-cell: LDS range of each of its 5 stencil planes
    __shared__ double red[NW];

    // LISTS: the last tile slot is a dummy far from every particle; lists
    // hold it in place of i itself and as padding to whole chunks, and the
    // cutoff test drops it
    const int kFar = T - 1;
    const int Tst = LISTS ? T - 1 : T;  // particles a pass can stage
    const Geom &g = P.g;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int grp = lane / kGL, gl = lane % kGL;
    const int bid = blk.x;
    const int bx = bid % g.nblk[0];
    const int by = (bid / g.nblk[0]) % g.nblk[1];
    const int bz = bid / (g.nblk[0] * g.nblk[1]);

    for (int t = tid; t < P.ntab * (D + 1); t += blockDim.x) {
        const int pc = t / (D + 1), k = t - pc * (D + 1);
        tab[pc * TS + k] = P.cheb[t];
    }
    if (LISTS && tid == 0) {
        if constexpr (kRel) {
            th4[kFar] = make_float4(1e18f, 1e18f, 1e18f, 0.0f);
            if constexpr (kPack) tl2[kFar] = make_float2(0.0f, 0.0f);
            else tl4[kFar] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            tx[kFar] = ty[kFar] = tz[kFar] = (R)1e150;
        }
    }
    // origin of the tile coordinates: 0, or the home block centre (mixed)
    const double ox = kRel ? (2 * bx + 1) * g.w[0] : 0.0;
    const double oy = kRel ? (2 * by + 1) * g.w[1] : 0.0;
    const double oz = kRel ? (2 * bz + 1) * g.w[2] : 0.0;
    // FP32: coordinates in units of 1 / sigma
    const double qs = kRel ? P.qscale : 1.0;
    // LDS address of the s-indexed table less its first base pieces (FP32)
    const unsigned tbase = (unsigned)(size_t)(__attribute__((address_space(3))) const R2 *)tab -
                           (kRel ? P.s_base * (kSStride * (unsigned)sizeof(float2)) : 0u);
    float s_top_v = kRel ? (float)P.s_top : 0.0f;
    asm volatile("" : "+v"(s_top_v));
    // 1. counts of the 216 tile sub-cells, tile order x fastest
    if (tid < 216) {
        const int txs = tid % 6, tys = (tid / 6) % 6, tzs = tid / 36;
        int cx, cy, cz;
        double sdummy;
        wrap_cell(2 * bx - 2 + txs, g.nsub[0], g.L[0], cx, sdummy);
        wrap_cell(2 * by - 2 + tys, g.nsub[1], g.L[1], cy, sdummy);
        wrap_cell(2 * bz - 2 + tzs, g.nsub[2], g.L[2], cz, sdummy);
        const int cell = cx + g.nsub[0] * (cy + g.nsub[1] * cz);
        const int s0 = P.cell_start[cell];
        tstart[tid] = s0;
        cum[tid + 1] = P.cell_start[cell + 1] - s0;
    }
    __syncthreads();
    // 2. exclusive prefix over the 216 counts (wave 0: 4 entries per lane)
    if (wave == 0) {
        int v[4];
        int s = 0;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
            const int k = lane * 4 + qq;
            v[qq] = k < 216 ? cum[k + 1] : 0;
            s += v[qq];
        }
        int incl = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        int run = incl - s;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
            const int k = lane * 4 + qq;
            if (k < 216) cum[k] = run;
            run += v[qq];
        }
        if (lane == 63) cum[216] = incl;
    }
    __syncthreads();

    // home block: tile sub-cells (2+hx, 2+hy, 2+hz), hx,hy,hz in {0,1}
    int hoff[9];
    hoff[0] = 0;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const int idx = (2 + (h & 1)) + 6 * ((2 + ((h >> 1) & 1)) + 6 * (2 + (h >> 2)));
        hoff[h + 1] = hoff[h] + (cum[idx + 1] - cum[idx]);
    }
    const int n_home = hoff[8];
    int *const qw = queue + __builtin_amdgcn_readfirstlane(wave) * (kGroups * kQCap);  // the wave's two queues
    int *q = qw + grp * kQCap;                                                           // this half's queue
    // all ones in the upper half; opaque to the compiler, which would
    // otherwise turn gmask & x into a select on grp (that needs x in a VGPR)
    unsigned gmask = grp ? ~0u : 0u;
    asm volatile("" : "+v"(gmask));
    // Verlet list mode of this step (uniform over the launch)
    const int vmode = LISTS ? __builtin_amdgcn_readfirstlane(P.vl_flags[3]) : 0;
    const bool build = vmode == 1, use_list = vmode == 2;
    const R rc2m = build ? P.rcl2 : P.rc2m;
    double u_acc = 0.0;

    int r0 = 0, pass = 0;
    while (r0 < 36) {
        const int base = cum[6 * r0];
        int r1 = r0;
        while (r1 < 36 && cum[6 * (r1 + 1)] - base <= Tst) ++r1;
        if (r1 == r0) {  // a single row larger than the LDS tile: unsupported density
            if (tid == 0) atomicOr(&P.stats[2], 1);
            r0 = r0 + 1;
            continue;
        }
        if (pass == 1 && tid == 0) {
            atomicAdd(&P.stats[0], 1);
            if (build) atomicOr(&P.vl_flags[1], 1);  // lists need the whole tile in one pass
        }
        // 3. stage rows [r0, r1) with their image shifts: each thread takes up
        // to three tile slots and issues their loads together (indices past the
        // pass clamped to its last slot, the loads unconditional), so a block
        // waits one or two memory latencies for its tile (four slots a thread
        // made the 12-wave FP32 kernel spill); row by row, a wave waited
        // one for each of its rows.  A slot's sub-cell comes from a binary
        // search of the sub-cell starts in LDS.
        {
            const int nst = cum[6 * r1] - base;
            constexpr int kSt = 3;
            for (int t0 = tid; t0 < nst; t0 += kSt * (int)blockDim.x) {
                double4 pv[kSt];
                int cv[kSt];
#pragma unroll
                for (int u = 0; u < kSt; ++u) {
                    const int x = min(t0 + u * (int)blockDim.x, nst - 1) + base;
                    int lo = 6 * r0, hi = 6 * r1;  // cum[lo] <= x < cum[hi]
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (cum[mid] <= x) lo = mid;
                        else hi = mid;
                    }
                    cv[u] = lo;
                    pv[u] = P.spos[tstart[lo] + (x - cum[lo])];
                }
#pragma unroll
                for (int u = 0; u < kSt; ++u) {
                    const int t = t0 + u * (int)blockDim.x;
                    if (t >= nst) continue;
                    const int c = cv[u], r = c / 6, txs = c - 6 * r, tys = r % 6, tzs = r / 6;
                    const double sx = image_shift(2 * bx - 2 + txs, g.nsub[0], g.L[0]);
                    const double sy = image_shift(2 * by - 2 + tys, g.nsub[1], g.L[1]);
                    const double sz = image_shift(2 * bz - 2 + tzs, g.nsub[2], g.L[2]);
                    const double4 p = pv[u];
                    if (kRel) {
                        const double cx = ((p.x + sx) - ox) * qs, cy = ((p.y + sy) - oy) * qs,
                                     cz = ((p.z + sz) - oz) * qs;
                        const float hx = (float)cx, hy = (float)cy, hz = (float)cz;
                        const float lx = (float)(cx - (double)hx), ly = (float)(cy - (double)hy),
                                    lz = (float)(cz - (double)hz);
                        if (kPack) {
                            th4[t] = make_float4(hx, hy, hz, lx);
                            tl2[t] = make_float2(ly, lz);
                        } else {
                            th4[t] = make_float4(hx, hy, hz, 0.0f);
                            tl4[t] = make_float4(lx, ly, lz, 0.0f);
                        }
                    } else {
                        tx[t] = (R)(p.x + sx);
                        ty[t] = (R)(p.y + sy);
                        tz[t] = (R)(p.z + sz);
                    }
                }
            }
        }
        // plane ranges of this pass: home sub-cell h needs tile x sub-cells
        // hx .. hx+4 of the five tile rows (hy+dy, hz+dz), dy = 0..4, of each
        // plane dz.  Those rows are consecutive in LDS, so one range from row
        // hy's sub-cell hx to row hy+4's sub-cell hx+4 covers them; it also
        // takes in one sub-cell three away along x between consecutive rows,
        // whose particles are at least 2 sub-cell widths >= rc away and fail
        // the cutoff test.  Rows outside this pass are left out.
        if (tid < 40) {
            const int h = tid / 5, pz = tid % 5;
            const int hx = h & 1, hy = (h >> 1) & 1, hz = h >> 2;
            const int ra = max(hy + 6 * (hz + pz), r0);
            const int rb = min(hy + 4 + 6 * (hz + pz), r1 - 1);
            int2 v = make_int2(0, 0);
            if (ra <= rb) v = make_int2(cum[6 * ra + hx] - base, cum[6 * rb + hx + 5] - base);
            rng[h][pz] = v;
        }
        __syncthreads();

        // 4. two i per wave, one per half
        for (int kb = wave * kGroups; kb < n_home; kb += NW * kGroups) {
            const int k = kb + grp;
            const bool have_i = k < n_home;
            const int kk = have_i ? k : kb;  // a valid home particle for the idle half
            int h = 0, hb = 0;
#pragma unroll
            for (int hh = 1; hh < 8; ++hh)
                if (kk >= hoff[hh]) {
                    h = hh;
                    hb = hoff[hh];
                }
            const int hx = h & 1, hy = (h >> 1) & 1, hz = h >> 2;
            const int hidx = (2 + hx) + 6 * ((2 + hy) + 6 * (2 + hz));
            const int gi = tstart[hidx] + (kk - hb);  // from LDS: no global load per i
            // tile slot of i in this pass (its row may be outside the pass)
            const int irow = (2 + hy) + 6 * (2 + hz);
            const int islot = (irow >= r0 && irow < r1) ? cum[hidx] - base + (kk - hb) : -1;
            // i's coordinates: from its tile slot when staged (home rows have image
            // shift 0, so the values are those of the staging), else from memory
            R xi, yi, zi, xil = 0, yil = 0, zil = 0;
            if (islot >= 0) {
                if constexpr (kRel) {
                    const float4 a = th4[islot];
                    xi = a.x;
                    yi = a.y;
                    zi = a.z;
                    if constexpr (kPack) {
                        const float2 b = tl2[islot];
                        xil = a.w;
                        yil = b.x;
                        zil = b.y;
                    } else {
                        const float4 b = tl4[islot];
                        xil = b.x;
                        yil = b.y;
                        zil = b.z;
                    }
                } else {
                    xi = tx[islot];
                    yi = ty[islot];
                    zi = tz[islot];
                }
            } else {
                const double4 pi = P.spos[gi];
                const double cxi = ((pi.x + 0.0) - ox) * qs, cyi = ((pi.y + 0.0) - oy) * qs,
                             czi = ((pi.z + 0.0) - oz) * qs;
                xi = (R)cxi;
                yi = (R)cyi;
                zi = (R)czi;
                if (kRel) {
                    xil = (R)(cxi - (double)xi);
                    yil = (R)(cyi - (double)yi);
                    zil = (R)(czi - (double)zi);
                }
            }
            R fx = 0, fy = 0, fz = 0, u = 0;
            auto disp = [&](int jj, R &dx, R &dy, R &dz) {
                if constexpr (kRel) {
                    const float4 a = th4[jj];
                    if constexpr (kPack) {
                        const float2 b = tl2[jj];
                        dx = (xi - a.x) + (xil - a.w);
                        dy = (yi - a.y) + (yil - b.x);
                        dz = (zi - a.z) + (zil - b.y);
                    } else {
                        const float4 b = tl4[jj];
                        dx = (xi - a.x) + (xil - b.x);
                        dy = (yi - a.y) + (yil - b.y);
                        dz = (zi - a.z) + (zil - b.z);
                    }
                } else {
                    dx = xi - tx[jj];
                    dy = yi - ty[jj];
                    dz = zi - tz[jj];
                }
            };
            auto pair_at = [&](int jj) {
                R dx, dy, dz;
                disp(jj, dx, dy, dz);
                if constexpr (kRel) pp_pair_f(P, tbase, s_top_v, dx, dy, dz, fx, fy, fz, u);
                else pp_pair_cheb<R>(P, tab, dx, dy, dz, fx, fy, fz, u);
            };
            int qn0 = 0, qn1 = 0;  // queue fill of half 0 / half 1 (wave-uniform)
            int lc0 = 0, lc1 = 0;  // list entries written for the i of half 0 / half 1 (build)
            unsigned short *const lst = build || use_list ? P.vl_list + (size_t)gi * P.vl_cap : nullptr;
            // build: the queued j of this half go to its i's list, slots lc .. lc + cnt - 1
            // (i itself, queued by the FP32 sweep, as the dummy slot)
            auto list_put = [&](int jj, int at) {
                if (at < P.vl_cap) lst[list_pos(at)] = (unsigned short)(jj == islot ? kFar : jj);
            };
            if (use_list) {
                // the step's pairs from i's list: every j within the list radius at
                // the last build, evaluated with the exact cutoff.  A wave runs
                // the half-chunks (kGL entries) of the longer of its two lists:
                // whole chunks (two entries a lane, one dword) and, for an odd
                // count, a last half-chunk (the low entries alone); padding to
                // whole half-chunks, not chunks, saves ~16 entries a list.  A
                // lane's chunks past its own list take the dummy pair (a select;
                // the load index is clamped to the list's last dword, the load
                // unconditional, its wait where the value is used)
                const int cnt = have_i ? P.vl_cnt[gi] : 0;
                const int hch = (cnt + kGL - 1) / kGL;  // this half's half-chunks
                const int hmax = max(readlane_i(hch, 0), readlane_i(hch, kGL));
                const int nchunk = hmax >> 1;      // whole chunks of the wave
                const bool tail = (hmax & 1) != 0;  // and a last half-chunk
                const int ownc = (hch + 1) >> 1;   // chunks of this half's list in memory
                const unsigned *lw = (const unsigned *)lst;  // vl_cap is a multiple of 64: dword aligned
                const int wlast = max(ownc * kGL - 1, 0);
                const unsigned kpad = (unsigned)kFar | ((unsigned)kFar << 16);
                auto chunk_word = [&](int c) -> unsigned { return lw[min(c * kGL + gl, wlast)]; };
                auto own_word = [&](int c, unsigned w) -> unsigned { return c < ownc ? w : kpad; };
                // FP32: the pair terms come from the table alone (s' clamped to
                // s_lo at the low end); close pairs (s' < s_lo, the release of
                // a random start) are corrected in a second pass over the list,
                // so the loop has no branch for them
                float smin = 3.0e38f;
                auto chunk_pairs = [&](unsigned w) {
                    const int j0 = (int)(w & 0xffffu), j1 = (int)(w >> 16);
                    R dx0, dy0, dz0, dx1, dy1, dz1;
                    disp(j0, dx0, dy0, dz0);
                    disp(j1, dx1, dy1, dz1);
                    if constexpr (kRel) {
                        float s0, s1;
                        pp_pair_tab(P, tbase, s_top_v, dx0, dy0, dz0, fx, fy, fz, u, s0);
                        pp_pair_tab(P, tbase, s_top_v, dx1, dy1, dz1, fx, fy, fz, u, s1);
                        smin = fminf(smin, fminf(s0, s1));
                    } else {
                        pp_pair_masked<R>(P, tab, true, dx0, dy0, dz0, fx, fy, fz, u);
                        pp_pair_masked<R>(P, tab, true, dx1, dy1, dz1, fx, fy, fz, u);
                    }
                };
                // the list words of the first KPF chunks are loaded up front, so
                // that their HBM latency overlaps the pair terms of the chunks
                // before them (a two-ahead register ring made the compiler wait
                // for each load at the next chunk: 0.773 -> 0.756 ms a list
                // step, N = 1e6 liquid); longer lists read the rest as they go
                constexpr int KPF = 8;
                unsigned wv[KPF];
#pragma unroll
                for (int c = 0; c < KPF; ++c) wv[c] = chunk_word(c);
#pragma unroll
                for (int c = 0; c < KPF; ++c)
                    if (c < nchunk) chunk_pairs(own_word(c, wv[c]));
                for (int c = KPF; c < nchunk; ++c) chunk_pairs(own_word(c, chunk_word(c)));
                if (tail) {  // the low entries of chunk nchunk
                    const unsigned w = own_word(nchunk, chunk_word(nchunk));
                    R dx, dy, dz;
                    disp((int)(w & 0xffffu), dx, dy, dz);
                    if constexpr (kRel) {
                        float s0;
                        pp_pair_tab(P, tbase, s_top_v, dx, dy, dz, fx, fy, fz, u, s0);
                        smin = fminf(smin, s0);
                    } else {
                        pp_pair_masked<R>(P, tab, true, dx, dy, dz, fx, fy, fz, u);
                    }
                }
                if constexpr (kRel) {
                    if (__builtin_amdgcn_ballot_w64(smin < P.s_lo)) {  // rare: close pairs in some list
                        for (int c = 0; c < nchunk + (tail ? 1 : 0); ++c) {
                            const unsigned w = own_word(c, chunk_word(c));
                            R dx, dy, dz;
                            disp((int)(w & 0xffffu), dx, dy, dz);
                            pp_pair_fix(P, tbase, s_top_v, dx, dy, dz, fx, fy, fz, u);
                            if (c < nchunk) {
                                disp((int)(w >> 16), dx, dy, dz);
                                pp_pair_fix(P, tbase, s_top_v, dx, dy, dz, fx, fy, fz, u);
                            }
                        }
                    }
                }
            }
            for (int rr = 0; rr < (use_list ? 0 : 5); ++rr) {
                int2 lr = rng[h][rr];
                if (!have_i) lr = make_int2(0, 0);
                const int lo = lr.x, hi = lr.y;
                const int len = hi - lo;
                const int nit = (max(readlane_i(len, 0), readlane_i(len, kGL)) + kGL - 1) / kGL;
                // one candidate batch: tile read and cutoff test, the hit mask in SGPRs
                auto test = [&](int j, bool &hit, unsigned long long &m) {
                    const bool inrow = j < hi;
                    const int jc = inrow ? j : 0;  // slot 0 is always inside the LDS tile
                    R ddx, ddy, ddz;
                    if constexpr (kRel) {
                        const float4 a = th4[jc];
                        ddx = xi - a.x;
                        ddy = yi - a.y;
                        ddz = zi - a.z;
                    } else {
                        ddx = xi - tx[jc];
                        ddy = yi - ty[jc];
                        ddz = zi - tz[jc];
                    }
                    const R r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    // FP32 has no self test here: i's own slot gives s' = 0
                    // exactly from the hi + lo parts, which the pair drops
                    const bool c1 = r2 < rc2m, c2 = kRel || r2 > (R)0;
                    hit = inrow & c1 & c2;
                    // the mask as the AND of the compares' own masks: no VALU to materialise it
                    m = __builtin_amdgcn_ballot_w64(inrow) & __builtin_amdgcn_ballot_w64(c1);
                    if constexpr (!kRel) m &= __builtin_amdgcn_ballot_w64(c2);
                };
                // queue the batch's hits; evaluate 32 pairs of a half whose queue filled
                auto enqueue = [&](int j, bool hit, unsigned long long m) {
                    const unsigned mlo = (unsigned)m, mhi = (unsigned)(m >> 32);
                    const int clo = __popc(mlo), chi = __popc(mhi);
                    // queue slot: the half's queue and fill plus the hits on lower
                    // lanes of the half.  mbcnt counts all of the lower half for
                    // upper lanes, so the upper half's slot is kQCap + qn1 - clo +
                    // mbcnt.  As qn0 + (gmask & qd) + mbcnt, with qd wave-uniform
                    // and qn0 folded into the wave's queue base, the per-half
                    // select is one AND (a select of two SGPR values would take
                    // two moves and a v_cndmask: one SGPR operand per VALU op)
                    // (the difference of the fills through readfirstlane: not
                    // reassociated with qn0 + clo, which would pull the fill
                    // update into the divergent branch below)
                    const unsigned qd = (unsigned)(__builtin_amdgcn_readfirstlane(kQCap + qn1 - qn0) - clo);
                    const int slot = (int)__builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, gmask & qd));
                    if (hit) qw[qn0 + slot] = j;
                    qn0 += clo;
                    qn1 += chi;
                    if (qn0 >= kGL || qn1 >= kGL) {
                        const int qmine = grp ? qn1 : qn0;
                        const int cnt = min(qmine, kGL);
                        if (gl < cnt) {
                            const int jj = q[gl];
                            if (build) list_put(jj, (grp ? lc1 : lc0) + gl);
                            pair_at(jj);
                        }
                        lc0 += min(qn0, kGL);
                        lc1 += min(qn1, kGL);
                        const int rest = qmine - cnt;
                        const int moved = gl < rest ? q[kGL + gl] : 0;
                        if (gl < rest) q[gl] = moved;
                        qn0 = max(qn0 - kGL, 0);
                        qn1 = max(qn1 - kGL, 0);
                    }
                };
                // two batches a trip, both tested before either is queued, so
                // their LDS reads and arithmetic overlap.  FP64 ends an odd count
                // with one batch (2% faster at rc 7); FP32 runs the last trip
                // whole, its second batch starting at lo + nit * kGL >= hi of both
                // halves with no hits (the peeled form measured 7% slower there)
                constexpr bool kPeel = sizeof(R) == 8;
                const int nit2 = kPeel ? (nit & ~1) : nit;
                int it = 0;
                for (; it < nit2; it += 2) {
                    const int j0 = lo + it * kGL + gl, j1 = j0 + kGL;
                    bool h0, h1;
                    unsigned long long m0, m1;
                    test(j0, h0, m0);
                    test(j1, h1, m1);
                    enqueue(j0, h0, m0);
                    enqueue(j1, h1, m1);
                }
                if (kPeel && it < nit) {
                    const int j0 = lo + it * kGL + gl;
                    bool h0;
                    unsigned long long m0;
                    test(j0, h0, m0);
                    enqueue(j0, h0, m0);
                }
            }
            const int qmine = grp ? qn1 : qn0;
            if (gl < qmine) {
                const int jj = q[gl];
                if (build) list_put(jj, (grp ? lc1 : lc0) + gl);
                pair_at(jj);
            }
            if (build && have_i) {
                const int total = (grp ? lc1 : lc0) + qmine;
                const int pend = list_padded(total);
                // dummy padding to the end of the last chunk (at most two slots a lane)
                if (pend <= P.vl_cap) {
                    for (int at = total + gl; at < pend; at += kGL) lst[list_pos(at)] = (unsigned short)kFar;
                }
                if (gl == 0) {
                    P.vl_cnt[gi] = min(total, P.vl_cap);
                    if (pend > P.vl_cap) atomicOr(&P.vl_flags[1], 1);
                }
            }
            double sfx, sfy, sfz, su;
            if constexpr (sizeof(R) == 4) {
                // FP32 sums of the lanes' FP32 partials (rounding ~1e-7 of the force)
                sfx = half_wave_sum(fx);
                sfy = half_wave_sum(fy);
                sfz = half_wave_sum(fz);
                su = half_wave_sum(u);
            } else {
                sfx = fx;
                sfy = fy;
                sfz = fz;
                su = u;
#pragma unroll
                for (int o = kGL / 2; o > 0; o >>= 1) {
                    sfx += __shfl_xor(sfx, o, 64);
                    sfy += __shfl_xor(sfy, o, 64);
                    sfz += __shfl_xor(sfz, o, 64);
                    su += __shfl_xor(su, o, 64);
                }
            }
            if (gl == 0 && have_i) {
                double *a = P.acc + 3 * (size_t)gi;
                if (pass == 0) {
                    a[0] = sfx;
                    a[1] = sfy;
                    a[2] = sfz;
                } else {
                    a[0] += sfx;
                    a[1] += sfy;
                    a[2] += sfz;
                }
                u_acc += su;
            }
        }
        __syncthreads();
        r0 = r1;
        ++pass;
    }
    u_acc += __shfl_xor(u_acc, kGL, 64);
    if (lane == 0) red[wave] = u_acc;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < NW; ++w) s += red[w];
        P.part[bid] = 0.5 * s;
    }
}

// The FP64 12-wave kernel asks for 6 waves a SIMD (80 VGPR, a 48-byte spill
// against 95 VGPR): two of its workgroups share a CU where one 16-wave
// workgroup did (rc 7 at unit density: PP 3.10 -> 2.49 ms).
// LISTS: the kernel of a context with pair lists (build or list mode, read
// from vl_flags[3] on the device); without, the list code is compiled out and
// the register count stays that of the sweep alone.
template <class R, int NW, bool LISTS>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(NW == 12 ? 6 : 1)))
void k_pp(PPArgs<R> P) {
    extern __shared__ __align__(16) unsigned char ldsraw[];
    pp_body<R, NW, LISTS>(P, blk_plain(), ldsraw);
}
// LISTS: a launch whose replicas all have pair lists (the segmented PP stage
// launches the replicas with and without lists apart).
template <class R, int NW, bool LISTS>
__global__ __launch_bounds__(NW * 64) void k_pp_seg(const PPArgs<R> *__restrict__ tab, SegGrid S) {
    extern __shared__ __align__(16) unsigned char ldsraw[];
    Blk b;
    const int r = seg_locate(S, b);
    pp_body<R, NW, LISTS>(tab[r], b, ldsraw);
}

size_t pp_lds_bytes(int tile_max, int npiece, bool mixed, int nw) {
    // tile: (x, y, z) in FP64 (24 B a particle); FP32 hi and lo float4 (32 B),
    // packed to 24 B in the 12-wave kernel
    const size_t es = mixed ? 4 : 8;
    const size_t ncoef = (size_t)npiece * (mixed ? kSStride : kChebDeg + 1);
    return ((mixed && nw != 12) ? 32 : 24) * (size_t)tile_max + 2 * es * ncoef + 4 * (size_t)(224 + nw * kGroups * kQCap);
}

template <class R>
static PPArgs<R> pp_args(mdf_p3m_ctx *c, const typename Vec2<R>::t *cheb, bool lists) {
    PPArgs<R> P;
    P.g = c->g;
    lists = lists && c->vl_skin > 0.0;
    P.spos = c->spos;
    P.cell_start = c->cell_start;
    P.vl_flags = lists ? c->vl_flags : nullptr;
    P.vl_list = c->vl_list;
    P.vl_cnt = c->vl_cnt;
    P.vl_cap = c->vl_cap;
    const double rl = (c->g.rc + c->vl_skin) * (1.0 + 1e-6);
    // FP32 compares squared distances in units of 1 / sigma^2
    const double q2 = sizeof(R) == 4 ? c->sp.qscale * c->sp.qscale : 1.0;
    P.rcl2 = (R)(rl * rl * q2);
    P.cheb = cheb;
    P.acc = c->acc_pp_s;
    P.part = c->part_pp;
    P.stats = c->stats;
    P.tile_max = c->tile_max;
    P.npiece = c->npiece;
    P.ntab = pp_table_pieces(c);
    P.piece_w = (R)(c->g.rc / c->npiece);
    P.inv_piece_w = (R)(c->npiece / c->g.rc);
    P.rtab = sizeof(R) == 4 ? (const typename Vec2<R>::t *)c->chebf : nullptr;
    P.s_base = c->sp.base;
    P.s_shift = c->sp.shift;
    P.s_lo = c->sp.s_lo;
    P.s_top = c->sp.s_top;
    P.s_tc = c->sp.tc;
    P.qscale = c->sp.qscale;
    P.qscale_f = (float)c->sp.qscale;
    P.inv_q = c->sp.inv_q;
    const double rcm = c->g.rc * (1.0 + 1e-12);
    // FP32: the cutoff is s' = s_top = sigma^2 rc^2 exactly (the table's zero piece)
    P.rc2 = sizeof(R) == 4 ? (R)c->sp.s_top : (R)(c->g.rc * c->g.rc);
    P.rc2m = sizeof(R) == 4 ? P.rc2 : (R)(rcm * rcm);
    return P;
}

template <class R, bool LISTS>
static int launch_pp_k(mdf_p3m_ctx *c, const PPArgs<R> &P, hipStream_t s) {
    if (c->pp_waves == 16) {
        MDF_TRY(allow_lds(k_pp<R, 16, LISTS>, c->pp_lds));
        hipLaunchKernelGGL((k_pp<R, 16, LISTS>), dim3(c->nblock), dim3(16 * 64), c->pp_lds, s, P);
    } else if (c->pp_waves == 12) {
        MDF_TRY(allow_lds(k_pp<R, 12, LISTS>, c->pp_lds));
        hipLaunchKernelGGL((k_pp<R, 12, LISTS>), dim3(c->nblock), dim3(12 * 64), c->pp_lds, s, P);
    } else if (c->pp_waves == 8) {
        MDF_TRY(allow_lds(k_pp<R, 8, LISTS>, c->pp_lds));
        hipLaunchKernelGGL((k_pp<R, 8, LISTS>), dim3(c->nblock), dim3(8 * 64), c->pp_lds, s, P);
    } else {
        MDF_TRY(allow_lds(k_pp<R, 4, LISTS>, c->pp_lds));
        hipLaunchKernelGGL((k_pp<R, 4, LISTS>), dim3(c->nblock), dim3(4 * 64), c->pp_lds, s, P);
    }
    MDF_LAUNCHED();
    return MDF_OK;
}

template <class R>
static int launch_pp(mdf_p3m_ctx *c, const typename Vec2<R>::t *cheb, hipStream_t s) {
    const PPArgs<R> P = pp_args<R>(c, cheb, true);
    if (P.vl_flags) return launch_pp_k<R, true>(c, P, s);
    return launch_pp_k<R, false>(c, P, s);
}

int pp_forces(mdf_p3m_ctx *c, int n, hipStream_t s) {
    (void)n;
    if (c->prm.precision == MDF_PREC_MIXED) return launch_pp<float>(c, c->stab, s);
    return launch_pp<double>(c, c->cheb, s);
}

// ---- segmented stage (replica batches) ----
template <class R>
static int build_pp(mdf_p3m_ctx *c, const StepIO &, void *args, int4 &dims, size_t &lds) {
    if ((sizeof(R) == 4) != (c->prm.precision == MDF_PREC_MIXED)) {
        set_error("replica batch: mixed precisions");
        return MDF_EINVAL;
    }
    const PPArgs<R> P = pp_args<R>(c, sizeof(R) == 4 ? (const typename Vec2<R>::t *)c->stab
                                                     : (const typename Vec2<R>::t *)c->cheb, true);
    *(PPArgs<R> *)args = P;
    dims = make_int4(c->nblock, 1, 1, P.vl_flags ? 1 : 0);  // variant 1: the list kernel
    // segmented launches run 16 waves for tiles above 64 KB, else 8 (FP32) or
    // 4 (FP64), whatever the replica's own launch uses; the LDS is sized for that
    const bool mixed = sizeof(R) == 4;
    const int nw = pp_lds_bytes(c->tile_max, pp_table_pieces(c), mixed, 4) > 64 * 1024 ? 16 : (mixed ? 8 : 4);
    lds = pp_lds_bytes(c->tile_max, pp_table_pieces(c), mixed, nw);
    return MDF_OK;
}
// A launch's LDS is the largest of its replicas' (each computed for the
// replica's segmented wave count, build_pp).  The per-wave scratch (the queues) is
// kGroups * kQCap * 4 = 512 B.  A replica runs 16 waves when its 4-wave size
// exceeds 64 KB, so its own size exceeds 70 KB, while 4- and 8-wave replicas
// stay within 64 + 2 KB; above 68 KB the group therefore holds a 16-wave
// replica, whose reservation also covers the 16-wave queues of any 4- or
// 8-wave replica beside it (at most 64 + 6 KB).
static_assert(kGroups * kQCap * 4 == 512, "per-wave PP scratch is 512 B");
template <class R, bool LISTS>
static int launch_pp_seg_k(const void *tab, const SegGrid &S, int total, size_t lds, hipStream_t s) {
    if (lds > 68 * 1024) {
        MDF_TRY(allow_lds(k_pp_seg<R, 16, LISTS>, lds));
        hipLaunchKernelGGL((k_pp_seg<R, 16, LISTS>), dim3(total), dim3(16 * 64), lds, s, (const PPArgs<R> *)tab, S);
    } else if (sizeof(R) == 4) {
        MDF_TRY(allow_lds(k_pp_seg<R, 8, LISTS>, lds));
        hipLaunchKernelGGL((k_pp_seg<R, 8, LISTS>), dim3(total), dim3(8 * 64), lds, s, (const PPArgs<R> *)tab, S);
    } else {
        MDF_TRY(allow_lds(k_pp_seg<R, 4, LISTS>, lds));
        hipLaunchKernelGGL((k_pp_seg<R, 4, LISTS>), dim3(total), dim3(4 * 64), lds, s, (const PPArgs<R> *)tab, S);
    }
    MDF_LAUNCHED();
    return MDF_OK;
}
template <class R>
static int launch_pp_seg(const void *tab, const SegGrid &S, int total, size_t lds, int, hipStream_t s) {
    if (S.variant == 1) return launch_pp_seg_k<R, true>(tab, S, total, lds, s);
    return launch_pp_seg_k<R, false>(tab, S, total, lds, s);
}

void seg_stages_pp(std::vector<SegStage> &out, int precision) {
    if (precision == MDF_PREC_MIXED)
        out.push_back({"pp", sizeof(PPArgs<float>), build_pp<float>, launch_pp_seg<float>});
    else
        out.push_back({"pp", sizeof(PPArgs<double>), build_pp<double>, launch_pp_seg<double>});
}

}  // namespace mdf
