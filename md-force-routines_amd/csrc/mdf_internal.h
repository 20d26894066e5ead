// Internal declarations shared by the mdforce HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/mdforce.h"

namespace mdf {

// ---------------------------------------------------------------- errors --
void set_error(const char *fmt, ...);

#define MDF_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::mdf::set_error("%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                \
            return MDF_EHIP;                                                           \
        }                                                                              \
    } while (0)

#define MDF_LAUNCHED()                                                                 \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess) {                                                        \
            ::mdf::set_error("kernel launch: %s (%s:%d)", hipGetErrorString(e_),       \
                             __FILE__, __LINE__);                                      \
            return MDF_EHIP;                                                           \
        }                                                                              \
    } while (0)

#define MDF_TRY(expr)            \
    do {                         \
        int r_ = (expr);         \
        if (r_ != MDF_OK) return r_; \
    } while (0)

// ------------------------------------------------------------- geometry --
// Box, mesh and cell geometry passed by value to kernels.
struct Geom {
    double L[3];      // box
    double h[3];      // mesh spacing L/M
    double ih[3];     // 1/h
    double w[3];      // PP sub-cell width L/nsub
    int M[3];         // mesh points
    int Mxh;          // M[0]/2 + 1 (half-spectrum length along x)
    int nsub[3];      // PP sub-cells per axis (even; two per reference cell)
    int nblk[3];      // PP blocks (2x2x2 sub-cells = one reference cell)
    double kappa, alpha, rc;
};

// Mesh bricks of the assignment and gather kernels.  Each workgroup handles
// a PM block of sb[0] x sb[1] x sb[2] sub-cells; its brick covers the
// stencils of the block's particles.
struct Brick {
    int B[3];          // assignment brick extent in mesh nodes
    int Bg[3];         // gather brick extent
    int sb_assign[3], sb_gather[3];  // sub-cells a block per axis
    int nblock_assign, nblock_gather;
    int lds_assign;    // 1 if the charge brick fits LDS for the assignment
    int lds_gather;    // 1 if the 3-field brick fits LDS for the gather
    int ord_cap;       // FP32 gather: particles a block the bank-spread visit order holds (0: none)
};

// FP32 gather brick (Ex, Ey, Ez a node): node (x, y, z) at float 3 x + Sy y +
// Sz z, Sy = 3 B0 and Sz the 3 B0 B1 floats of a plane rounded up to 16 (mod
// 32).  The gather's LDS reads behave as 32 banks of 4 bytes served a
// half-wave at a time (its bank-conflict counters follow that model to within
// 5%); a particle's mirrored plane pairs (offsets 5, 1 and -3 planes) are then
// 16 banks apart and its two x halves (nodes 0-2 and 3-5) 9 banks apart, so
// the four lanes of a particle hit four banks whose set is fixed by its base
// address mod 16.  The kernel visits a block's particles in an order that puts
// 8 particles of distinct base (mod 16, one parity) in each half-wave, which
// makes a half-wave's reads conflict-free.
__host__ __device__ inline int gather_plane_f32(int B0, int B1) {
    const int s = 3 * B0 * B1;
    return s + ((16 - s) & 31);
}
// dynamic LDS of the gather: the field brick, plus (FP32) the visit order
inline size_t gather_brick_bytes(const int *Bg, bool mixed) {
    return mixed ? 4 * (size_t)gather_plane_f32(Bg[0], Bg[1]) * Bg[2] : 24 * (size_t)Bg[0] * Bg[1] * Bg[2];
}
inline size_t gather_lds_bytes(const Brick &b, bool mixed) {
    return gather_brick_bytes(b.Bg, mixed) + (mixed ? 2 * (size_t)b.ord_cap : 0);
}

// FP32 pair table indexed by the float bits of s' = sigma^2 r^2, the squared
// distance in the mixed PP kernel's tile units 1 / sigma (pair_table.hip): piece
// p = (bits(s') >> shift) - base for s_lo <= s' <= s_top = 2^E = sigma^2 rc^2,
// the last piece (at s_top) all zeros; the local variable is f - tc, with f
// the float of s''s low `shift` mantissa bits under the exponent of 1 and tc
// the centre of f's range.
struct SPieces {
    int npiece;      // pieces including the zero piece at s_top
    unsigned base;
    int shift;
    float s_lo, s_top, tc;
    int E;           // s_top = 2^E
    double qscale;   // sigma: tile units per length unit
    float inv_q;     // 1 / sigma
};

constexpr int kPMThreads = 512;
constexpr int kGatherThreads = 1024;
constexpr int kChebDeg = 13;        // degree of the PP pair-term Chebyshev pieces (FP64)
constexpr int kChebDegF = 4;        // degree of the FP32 r-indexed pieces (mixed precision)
// degree of the FP32 s-indexed pieces (the pair table in LDS).  Degree 3
// (238 pieces at kappa 2, alpha 0.43, rc 6.75, against 119) measured 1.450
// against 1.430 ms/step at N = 1e6: one LDS read and two FMAs fewer a pair,
// but the 3 KB larger table cost the two 12-wave workgroups a CU at rc 7
constexpr int kSDegF = 4;
// LDS stride of an s-indexed piece in (phi, F) coefficient pairs: an odd
// count, so that a piece spans 2 (mod 4) dwords and the pieces of a wave's
// lanes spread over the 64 LDS banks (at 8 dwords, pieces 8 apart share banks:
// degree 3 unpadded ran 1.74 against 1.45 ms/step padded)
constexpr int kSStride = (kSDegF + 1) | 1;
constexpr int kMaxPartials = 1 << 20;

// ctx->stats (device ints): [0] PP tiles needing more than one LDS pass,
// [1] particles outside their assignment/gather brick (handled through
// global memory), [2] error flag: a tile row exceeded the LDS capacity.

}  // namespace mdf

struct mdf_p3m_ctx {
    mdf_p3m_params prm;
    mdf::Geom g;
    mdf::Brick brick;
    int ncell;          // total sub-cells
    int nblock;         // total PP blocks
    int tile_max;       // LDS tile capacity (particles)
    size_t pp_lds;      // dynamic LDS bytes of the PP kernel
    int pp_waves;       // waves per PP workgroup (4, 8, 12 or 16)
    int shape_assign;   // PM block shapes in use (1-based index into the shape list of create)
    int shape_gather;
    // launch choices asked for by mdf_p3m_set_launch (0 = chosen at create) and
    // the skin asked for by mdf_p3m_set_skin; kept when the context is rebuilt
    int req_waves, req_assign, req_gather;
    double skin_req;
    size_t nmesh;       // Mx*My*Mz
    size_t nspec;       // Mz*My*Mxh
    bool have_green;

    // binning
    int *cell_count;    // ncell, zero between steps
    int *cell_start;    // ncell + 1
    int *scan_tmp;      // per-tile sums of the count scan (+1 total)
    int *part_cell;     // n_max
    int *part_slot;     // n_max
    int *perm;          // sorted -> caller index
    double4 *spos;      // sorted (x, y, z, Z), positions wrapped into [0, L]
    double *acc_pp_s;   // sorted PP acceleration (n, 3)
    double2 *cheb;      // PP pair-term tables, npiece * (kChebDeg + 1) (g1, g2) pairs
    int npiece;
    float2 *chebf;      // FP32 tables of the mixed mode, npiece * (kChebDegF + 1)
    float2 *stab;       // FP32 s-indexed pair table of the mixed mode, sp.npiece * (kSDegF + 1)
    mdf::SPieces sp;

    // mesh
    double *rho;        // nmesh
    double2 *ck;        // nspec
    double2 *ek;        // 3 * nspec
    double *emesh;      // 3 * nmesh, interleaved (Ex, Ey, Ez) per node
    double *green;      // nspec, natural order half spectrum
    double2 *tw[3];     // twiddles per axis, M/2 entries
    // FP32 copies of the mixed mode (MDF_PREC_MIXED); the FP64 ck/ek/emesh are
    // then not allocated
    float2 *ckf, *ekf;
    float *emeshf;
    float *greenf;
    float2 *twf[3];

    // reductions
    double *part_pp;    // per-block PP energy partials (nblock)
    double *part_pm;    // per-block PM energy partials
    int n_part_pm;
    double *part_red;   // generic partials (kMaxPartials doubles)
    double *scal;       // device scalars, see kScal*
    int *stats;         // device Stats
    struct Prof *prof;  // optional stage timing (null when off)
    mdf_mlp *mlp;       // attached learned PP correction (null = none)

    // Verlet pair lists of the PP stage (vl_skin > 0: on; mdf_p3m_set_skin).
    // Lists hold, per particle, the tile slots of every j within rc + skin at
    // the last build, in the sorted order and cell starts of that build, which
    // the binning then keeps (frozen) until the next build; they stay valid
    // while no particle has moved more than skin / 2.  On a list step spos
    // holds the step's positions in the frozen order.
    double vl_skin;
    double list_skin_req;    // cap on vl_skin (mdf_p3m_set_list_skin; 0: the largest the cells allow)
    int vl_cap;              // list entries per particle
    unsigned short *vl_list; // n_max * vl_cap
    int *vl_cnt;             // n_max
    double4 *vl_ref;         // frozen order: positions at the build
    int *vl_flags;           // [0] build requested, [1] last build incomplete, [3] mode of the step
                             // (1 build, 2 lists), [4] builds, [5] list steps (counts)
    int vl_force_build;      // host: the next list-mode step must build
    int vl_n;                // host: particle count of the last step prepared (a change builds)
    int gen;                 // host: in-place rebuilds so far (captured graphs hold its buffers)
};

namespace mdf {

// Device scalar slots in ctx->scal.
enum {
    kScalU = 0, kScalUpp, kScalUpm, kScalUself,  // energies (U first, as returned)
    kScalK, kScalT, kScalFact, kScalPad,
    kScalCount = 16
};

// Stage ids of the optional event timing (mdf_p3m_profile).
enum {
    kStageBin = 0, kStagePP, kStageAssign, kStageFFT, kStageGather, kStageDrift, kStageModes,
    kStageFinal, kStageCount, kStageEnd = -1
};
// Records an event on s opening `stage` (or closing the previous one with
// kStageEnd); a no-op unless profiling is on.
void prof_mark(mdf_p3m_ctx *c, int stage, hipStream_t s);

// ----------------------------------------------------------- launchers --
// Counting sort of the positions into the sorted copy.  bin_particles: always
// (with pair lists on, the next list-aware step then builds);
// bin_particles_mode: skipped on the device when *mode == 2 (a list step).
int bin_particles(mdf_p3m_ctx *c, const double *pos, const double *Z, int n, hipStream_t s);
// vl_ref: the pair lists' build positions, written with the sorted copy (or null)
int bin_particles_mode(mdf_p3m_ctx *c, const double *pos, const double *Z, int n, const int *mode, hipStream_t s,
                       double4 *vl_ref = nullptr);
int pp_forces(mdf_p3m_ctx *c, int n, hipStream_t s);
int pm_assign(mdf_p3m_ctx *c, int n, double *rho_out, hipStream_t s);
int pm_solve(mdf_p3m_ctx *c, hipStream_t s);
// Gather plus epilogue: acc (caller order) = af*(acc_pp_s + Z*E_p); optional
// mirrors; if kick_dt != 0 also vel += 0.5*acc*kick_dt with K partials.
struct GatherOut {
    double *acc;      // (n,3) or null
    double af;
    int include_pp;   // add acc_pp_s into acc
    double *acc_pp;   // mirror (n,3) or null
    double *acc_pm;   // mirror (n,3) or null
    double *ex, *ey, *ez;
    double *vel;      // for the fused second half-kick, or null
    double kick_dt;
};
int pm_gather(mdf_p3m_ctx *c, int n, const GatherOut &o, hipStream_t s);
int finalize_energy(mdf_p3m_ctx *c, int n, double uf, int have_pp, int have_pm, double *energies,
                    hipStream_t s);
int build_green(mdf_p3m_ctx *c, int p, int mx, int my, int mz, double *G_full, double *A_pm,
                hipStream_t s);
int pm_error_sum(mdf_p3m_ctx *c, int p, int mx, int my, int mz, double *Q_sum, hipStream_t s);
int set_green(mdf_p3m_ctx *c, const double *G_full, hipStream_t s);

// PP pair-term tables (host) and the PP kernel's dynamic LDS size.
int build_pair_table(double kappa, double alpha, double rc, int deg, double tol, bool as_float,
                     std::vector<double2> &tab, int &npiece);
int build_pair_table_s(double kappa, double alpha, double rc, int deg, double tol, int E, std::vector<double2> &tab,
                       SPieces &sp);
size_t pp_lds_bytes(int tile_max, int npiece, bool mixed, int nw);
// pieces of the PP table staged in LDS: the s-indexed table (FP32) or the
// r-indexed Chebyshev table (FP64)
inline int pp_table_pieces(const mdf_p3m_ctx *c);
// FP32 copy of the FP64 half-spectrum Green function (mixed mode).
int green_to_float(mdf_p3m_ctx *c, hipStream_t s);
// Learned correction of the attached network, added to acc_pp_s.
int mlp_correct(mdf_mlp *m, mdf_p3m_ctx *c, int n, hipStream_t s);
// The network's pair-list overflow flag (device int).
int *mlp_flag(mdf_mlp *m);
// Overflow flags that the kernels set where they would otherwise leave work
// undone: a PP tile row over the LDS capacity (ctx stats[2]) and, when m is
// not null, a pair list over the network's capacity.  Waits for the stream's
// work, returns MDF_ESTATE with a message if either is set and clears both.
// A no-op on a stream that is being captured into a graph.
int check_overflow(mdf_p3m_ctx *c, mdf_mlp *m, hipStream_t s);

// Fill helpers (kernels, so they can live inside captured graphs).
int fill_zero(double *p, size_t n, hipStream_t s);
int fill_zero_i(int *p, size_t n, hipStream_t s);

// Integrator pieces.
int kick_drift(mdf_p3m_ctx *c, double *pos, double *vel, const double *acc, int n, double dt,
               int pbc, const double lmin[3], const double lmax[3], const double *scale,
               hipStream_t s);
int kick(mdf_p3m_ctx *c, double *vel, const double *acc, int n, double dt, double *k_sum,
         hipStream_t s);
int density_modes(mdf_p3m_ctx *c, const double *pos, int n, const double *q, int nq,
                  double *out, hipStream_t s);
int density_modes_grid(mdf_p3m_ctx *c, const double *pos, int n, double dq, int nq, double *out,
                       hipStream_t s);
// Copy a sorted (n,3) array back to the caller's particle order.
int unpermute3(mdf_p3m_ctx *c, const double *src_sorted, double *dst, int n, hipStream_t s);

// The binning of a force step (vlist.hip).  Without pair lists: the counting
// sort.  With them: the positions into the frozen order and the decision
// whether the lists still hold (no particle moved more than skin / 2 since the
// build, no incomplete build, no build requested), which sets the mode the
// binning and the PP kernel read on the device; a build step bins anew and
// records the build positions.  pos: caller order.
int bin_step(mdf_p3m_ctx *c, const double *pos, const double *Z, int n, hipStream_t s);
int vlist_alloc(mdf_p3m_ctx *c, double skin);
// Requests a build at the next list-aware step from the stream (outside a
// captured graph: the particles of n may have moved since the last step).
int vlist_request_build(mdf_p3m_ctx *c, int n, hipStream_t s);
void vlist_free(mdf_p3m_ctx *c);

// Dynamic LDS above 64 KiB (gfx950 has 160 KiB per workgroup) must be
// allowed per kernel before launch.
template <class F>
inline int allow_lds(F *kernel, size_t bytes) {
    if (bytes <= 65536) return MDF_OK;
    hipError_t e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // do not leave the error to the next launch check
        set_error("hipFuncSetAttribute(MaxDynamicSharedMemorySize=%zu): %s", bytes, hipGetErrorString(e));
        return MDF_EHIP;
    }
    return MDF_OK;
}

// Block coordinates seen by a kernel body.  A plain launch passes its own
// blockIdx/gridDim; a segmented launch (one launch for every replica of a
// batch) passes the block's place inside its replica's grid.
struct Blk {
    int x, y, z, nx, ny, nz;
};
__device__ __forceinline__ Blk blk_plain() {
    return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}
// Segmented grids: replica r owns the 1-D blocks [first[r], first[r+1]) and
// its own grid dims[r] (flattened x fastest).
struct SegGrid {
    const int *first;
    const int4 *dims;  // (nx, ny, nz, variant)
    int nrep;
    int variant;       // host: the kernel variant of the launch (dims[r].w of its replicas)
};
__device__ __forceinline__ int seg_locate(const SegGrid &S, Blk &b) {
    const int g = (int)blockIdx.x;
    int lo = 0, hi = S.nrep - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.first[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const int l = g - S.first[lo];
    const int4 d = S.dims[lo];
    b = Blk{l % d.x, (l / d.x) % d.y, l / (d.x * d.y), d.x, d.y, d.z};
    return lo;
}

// Segmented form of the fused NVE step (replica batches).
// Per-replica inputs of one step; per-step records of step k go to
// rec0 + k * rec_stride and modes0 + k * modes_stride.
struct StepIO {
    double *pos, *vel, *acc;
    int n;
    double dt, af, uf, kf;
    double *rec0;
    int rec_stride;
    double *modes0;
    int modes_stride;
    int nq;
};
// One stage: build() fills the replica's argument block, its grid and its
// dynamic LDS bytes; launch() runs the stage for all replicas at once.
struct SegStage {
    const char *name;
    size_t arg_size;
    int (*build)(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds);
    int (*launch)(const void *tab, const SegGrid &S, int total, size_t lds, int k, hipStream_t s);
};
// Stages of md_step followed by the n(q) grid modes, in order, for one precision.
int seg_step_stages(int precision, std::vector<SegStage> &out);
void seg_stages_md(std::vector<SegStage> &out, int part);    // part 0: drift; 1: finalize; 2: modes
void seg_stages_bin(std::vector<SegStage> &out);
void seg_stages_pp(std::vector<SegStage> &out, int precision);
void seg_stages_pm(std::vector<SegStage> &out, int precision, int part);  // 0: fill + assign, 1: gather
void seg_stages_fft(std::vector<SegStage> &out, int precision);
void seg_stages_vl(std::vector<SegStage> &out, int part);  // 0: prepare + decide, 1: refresh

inline int pp_table_pieces(const mdf_p3m_ctx *c) {
    return c->prm.precision == MDF_PREC_MIXED ? c->sp.npiece : c->npiece;
}

// Wave (64-lane) helpers.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace mdf
