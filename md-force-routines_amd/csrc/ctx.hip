// Context management and the extern "C" entry points of libmdforce.
//
// The entry points replace the reference's Python force API
// (yukawa_p3m_ord6.force_pot, yukawa_p3m_ord6.py:6-29) and its children; the
// Python layer (mdforce.yukawa_p3m_ord6 and friends) keeps the reference's
// positional signatures and calls these through ctypes.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "mdf_internal.h"

namespace mdf {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int md_step(mdf_p3m_ctx *c, double *pos, double *vel, double *acc, const double *Z, int n, double dt, double af,
            double uf, double kf, double *rec, hipStream_t s);
int thermostat_scale(mdf_p3m_ctx *c, double *vel, int n, const double *k_sum, double T_desired, int it,
                     double *t_out, hipStream_t s);

static bool pow2(int m) { return m > 0 && (m & (m - 1)) == 0; }

}  // namespace mdf

// Stage timing: a pool of events recorded in stream order; the interval
// between consecutive marks is charged to the stage of the first mark.
struct Prof {
    std::vector<hipEvent_t> ev;
    std::vector<int> stage;
    size_t used = 0;
    bool overflow = false;
    bool pp_only = false;  // marks at the PP stage's start and end only (two events a step)
};

namespace mdf {

void prof_mark(mdf_p3m_ctx *c, int stage, hipStream_t s) {
    Prof *p = c->prof;
    if (!p) return;
    if (p->pp_only && stage != kStagePP && stage != kStageAssign) return;
    if (p->used >= p->ev.size()) {
        p->overflow = true;
        return;
    }
    (void)hipEventRecord(p->ev[p->used], s);
    p->stage[p->used] = stage;
    ++p->used;
}

static void prof_free(mdf_p3m_ctx *c) {
    if (!c->prof) return;
    for (hipEvent_t e : c->prof->ev) (void)hipEventDestroy(e);
    delete c->prof;
    c->prof = nullptr;
}

template <class T>
static int dmalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        *p = nullptr;
        return MDF_ENOMEM;
    }
    return MDF_OK;
}

static void free_ctx(mdf_p3m_ctx *c) {
    if (!c) return;
    prof_free(c);
    void *ptrs[] = {c->cheb,     c->chebf,     c->scan_tmp, c->cell_count, c->cell_start, c->part_cell,
                    c->part_slot, c->perm,      c->spos,     c->acc_pp_s,   c->rho,        c->ck,
                    c->ek,        c->emesh,     c->green,    c->tw[0],      c->tw[1],      c->tw[2],
                    c->ckf,       c->ekf,       c->emeshf,   c->greenf,     c->twf[0],     c->twf[1],
                    c->twf[2],    c->part_pp,   c->part_pm,  c->part_red,   c->scal,       c->stats};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->stab) (void)hipFree(c->stab);
    vlist_free(c);
    delete c;
}

static int check_n(mdf_p3m_ctx *c, int n) {
    if (n < 0 || n > c->prm.n_max) {
        set_error("particle count %d outside [0, n_max=%d]", n, c->prm.n_max);
        return MDF_EINVAL;
    }
    return MDF_OK;
}

static int need_green(mdf_p3m_ctx *c) {
    if (!c->have_green) {
        set_error("no Green function: call mdf_p3m_set_green or mdf_p3m_build_green first");
        return MDF_ESTATE;
    }
    return MDF_OK;
}

int check_overflow(mdf_p3m_ctx *c, mdf_mlp *m, hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MDF_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return MDF_OK;
    int h[4] = {0, 0, 0, 0}, hm = 0;
    MDF_HIP(hipMemcpyAsync(h, c->stats, sizeof h, hipMemcpyDeviceToHost, s));
    if (m) MDF_HIP(hipMemcpyAsync(&hm, mlp_flag(m), sizeof hm, hipMemcpyDeviceToHost, s));
    MDF_HIP(hipStreamSynchronize(s));
    if (!h[2] && !hm) return MDF_OK;
    if (h[2]) MDF_TRY(fill_zero_i(c->stats + 2, 1, s));
    if (hm) MDF_TRY(fill_zero_i(mlp_flag(m), 1, s));
    MDF_HIP(hipStreamSynchronize(s));
    if (h[2])
        set_error("a PP tile row held more particles than the LDS tile capacity (%d): its pairs were not "
                  "evaluated; the density is beyond what this context was sized for (raise n_max)",
                  c->tile_max);
    else
        set_error("the learned correction's pair list overflowed its capacity: its pairs were not all evaluated");
    return MDF_ESTATE;
}

}  // namespace mdf

using namespace mdf;

extern "C" {

int mdf_abi_version(void) { return MDF_ABI_VERSION; }

const char *mdf_last_error(void) { return g_err; }

}  // extern "C"

namespace mdf {

// A context whose cell grid is sized for rc + skin (skin = 0: the reference's
// cells of L / floor(L / rc), yukawa_pp.py:17-23); the pair terms and the
// cutoff stay rc.  skin > 0 also allocates the Verlet pair lists.
// req: PP waves per workgroup, assignment and gather block shapes (1-based
// indices into the shape list below); 0 leaves the choice to this function.
static double list_skin_of(const mdf_p3m_ctx *c, double skin, double cap);

static int create_ctx(const mdf_p3m_params *p, double skin, const int req[3], double list_cap, mdf_p3m_ctx **out) {
    if (!p || !out) {
        set_error("mdf_p3m_create: null argument");
        return MDF_EINVAL;
    }
    *out = nullptr;
    if (!(skin >= 0.0) || !(skin < 1e300)) {
        set_error("pair-list skin %g: need a finite skin >= 0", skin);
        return MDF_EINVAL;
    }
    if (req[0] != 0 && req[0] != 4 && req[0] != 8 && req[0] != 12 && req[0] != 16) {
        set_error("PP waves per workgroup %d: need 4, 8, 12 or 16 (0 = automatic)", req[0]);
        return MDF_EINVAL;
    }
    if (req[1] < 0 || req[1] > 4 || req[2] < 0 || req[2] > 10) {
        set_error("PM block shapes %d / %d: need 0..4 (assignment) and 0..10 (gather), 0 = automatic", req[1],
                  req[2]);
        return MDF_EINVAL;
    }
    const double rcg = p->rc + skin;  // the cells hold every pair within rc + skin
    if (p->order != 6) {
        set_error("charge-assignment order %d not supported (only 6, as yukawa_pm_ord6)", p->order);
        return MDF_EINVAL;
    }
    if (p->precision != MDF_PREC_FP64 && p->precision != MDF_PREC_MIXED) {
        set_error("precision mode %d unknown (MDF_PREC_FP64 = 0, MDF_PREC_MIXED = 1)", p->precision);
        return MDF_EINVAL;
    }
    if (p->n_max <= 0) {
        set_error("n_max must be positive");
        return MDF_EINVAL;
    }
    if (!(p->alpha > 0.0) || !(p->rc > 0.0) || !(p->kappa >= 0.0)) {
        set_error("need alpha > 0, rc > 0, kappa >= 0");
        return MDF_EINVAL;
    }
    for (int d = 0; d < 3; ++d) {
        if (!pow2(p->mesh[d]) || p->mesh[d] < 8 || p->mesh[d] > 512) {
            set_error("mesh[%d] = %d: need a power of two in [8, 512]", d, p->mesh[d]);
            return MDF_EINVAL;
        }
        if (!(p->box[d] > 0.0)) {
            set_error("box[%d] must be positive", d);
            return MDF_EINVAL;
        }
        if ((int)floor(p->box[d] / rcg) < 3) {
            if (skin > 0.0)
                set_error("rc + skin = %g too large for box %g: the cell list needs L/(rc + skin) >= 3 per axis", rcg,
                          p->box[d]);
            else
                set_error("rc = %g too large for box %g: the cell list needs L/rc >= 3 per axis", p->rc, p->box[d]);
            return MDF_EINVAL;
        }
    }
    {
        // cell counts (reference cells floor(L/rc) per axis, two sub-cells each),
        // refused before anything is allocated
        double nb = 1.0;
        for (int d = 0; d < 3; ++d) nb *= floor(p->box[d] / rcg);
        const double ns = 8.0 * nb;
        if (ns > 1024.0 * 8192.0) {
            set_error("too many cells (%.0f sub-cells > %d): increase rc", ns, 1024 * 8192);
            return MDF_EINVAL;
        }
        if (nb > (double)kMaxPartials) {
            set_error("too many cells (%.0f > %d): increase rc", nb, kMaxPartials);
            return MDF_EINVAL;
        }
    }
    mdf_p3m_ctx *c = new (std::nothrow) mdf_p3m_ctx();
    if (!c) {
        set_error("out of host memory");
        return MDF_ENOMEM;
    }
    std::memset((void *)c, 0, sizeof(*c));
    c->prm = *p;
    Geom &g = c->g;
    for (int d = 0; d < 3; ++d) {
        g.L[d] = p->box[d];
        g.M[d] = p->mesh[d];
        g.h[d] = p->box[d] / p->mesh[d];
        g.ih[d] = 1.0 / g.h[d];
        const int lxd = (int)floor(p->box[d] / rcg);  // reference cells, yukawa_pp.py:17-19
        g.nsub[d] = 2 * lxd;
        g.nblk[d] = lxd;
        g.w[d] = p->box[d] / g.nsub[d];
    }
    g.Mxh = g.M[0] / 2 + 1;
    g.kappa = p->kappa;
    g.alpha = p->alpha;
    g.rc = p->rc;
    c->ncell = g.nsub[0] * g.nsub[1] * g.nsub[2];
    c->nblock = g.nblk[0] * g.nblk[1] * g.nblk[2];
    const bool mixed = p->precision == MDF_PREC_MIXED;
    const size_t es = mixed ? 4 : 8;  // bytes of the pair/mesh arithmetic type
    c->nmesh = (size_t)g.M[0] * g.M[1] * g.M[2];
    c->nspec = (size_t)g.M[2] * g.M[1] * g.Mxh;

    // LDS tile capacity of the PP kernel: mean tile + 6 sigma (Poisson) + slack;
    // larger tiles take more than one pass
    // (with pair lists 4 sigma + 32: a build whose tile overflows only builds
    // again on the next step, and the smaller tile keeps two PP workgroups on
    // a CU at the larger cells of rc + skin)
    const double mean_tile = 216.0 * (double)p->n_max / (double)c->ncell;
    double tm = skin > 0.0 ? mean_tile + 4.0 * sqrt(mean_tile) + 32.0 : mean_tile + 6.0 * sqrt(mean_tile) + 64.0;
    const double tcap = (150.0 * 1024.0 - 5120.0) / (mixed ? 32.0 : 24.0);
    if (tm > tcap) tm = tcap;
    if (tm < 512.0) tm = 512.0;
    c->tile_max = ((int)tm + 3) & ~3;  // keeps the table after the tile 16-byte aligned
    std::vector<double2> pair_tab, s_tab;
    {
        int rr = mixed ? build_pair_table(p->kappa, p->alpha, p->rc, kChebDegF, 1e-7, true, pair_tab, c->npiece)
                       : build_pair_table(p->kappa, p->alpha, p->rc, kChebDeg, 4e-15, false, pair_tab, c->npiece);
        // unit of the mixed PP kernel's tile coordinates: 1 / sigma with
        // sigma = 2^(E/2) / rc in [1, sqrt 2), so that the cutoff is the
        // squared distance 2^E
        const int E = (int)ceil(2.0 * log2(p->rc));
        if (rr == MDF_OK && mixed) rr = build_pair_table_s(p->kappa, p->alpha, p->rc, kSDegF, 1e-7, E, s_tab, c->sp);
        if (rr != MDF_OK) {
            delete c;
            return rr;
        }
    }
    // waves per PP workgroup.  FP64: below.  FP32 (71 VGPR; counted as at most 6 waves a
    // SIMD = 24 a CU: a 14-wave kernel aiming at 28 ran 1.5x slower, as if
    // its 4,4,3,3 split over the SIMDs kept a second workgroup out): the
    // count of 8, 12 or 16 that puts the most waves on a CU given the
    // LDS of each workgroup (dynamic + ~1 KB static of 160 KB), the fewest
    // waves on a tie; 16 when even one workgroup's tile overflows the LDS
    // (multipass).  The 12-wave kernel packs its tile to 24 B a particle
    // (32 B otherwise): at rc = 2w and unit density (2568-particle tiles) that
    // is two workgroups of 12 waves, against one of 16.
    if (req[0]) {
        c->pp_waves = req[0];
    } else if (mixed) {
        int best = 16, best_waves = 0;
        for (int nw : {8, 12, 16}) {
            const size_t l = pp_lds_bytes(c->tile_max, pp_table_pieces(c), true, nw) + 1024;
            const int wg = (int)std::min<size_t>((160 * 1024) / l, (size_t)(24 / nw));
            if (wg * nw > best_waves) {
                best_waves = wg * nw;
                best = nw;
            }
        }
        c->pp_waves = best;
    } else {
        // FP64: 4 waves for tiles up to 64 KB (several workgroups a CU); above
        // that two workgroups of 12 (80 VGPR, see k_pp) when their LDS fits,
        // else one of 16.  At rc 5.5 the 4-wave layout beat 12 waves with
        // the spill (1.69 against 1.78 ms), so 12 replaces 16 only.
        const bool big = pp_lds_bytes(c->tile_max, pp_table_pieces(c), false, 4) > 64 * 1024;
        const bool two12 = 2 * (pp_lds_bytes(c->tile_max, pp_table_pieces(c), false, 12) + 1536) <= 160 * 1024;
        c->pp_waves = !big ? 4 : (two12 ? 12 : 16);
    }
    c->pp_lds = pp_lds_bytes(c->tile_max, pp_table_pieces(c), mixed, c->pp_waves);
    // dynamic LDS plus the kernel's static arrays (< 4 KB) must fit 160 KB, for
    // this wave count and for that of segmented launches (16 above 64 KB)
    const size_t lds_cap = 160 * 1024 - 4096;
    auto lds_need = [&]() {
        return std::max(pp_lds_bytes(c->tile_max, pp_table_pieces(c), mixed, c->pp_waves),
                        pp_lds_bytes(c->tile_max, pp_table_pieces(c), mixed, 16));
    };
    while (lds_need() > lds_cap && c->tile_max > 0)
        c->tile_max -= (int)(((lds_need() - lds_cap) / 32 + 4) & ~(size_t)3);
    c->pp_lds = pp_lds_bytes(c->tile_max, pp_table_pieces(c), mixed, c->pp_waves);
    if (c->tile_max < 256) {
        delete c;
        set_error("PP pair table too large for LDS (kappa=%g alpha=%g rc=%g)", p->kappa, p->alpha, p->rc);
        return MDF_EINVAL;
    }
    if (req[0] && c->tile_max < ((int)(0.8 * mean_tile) & ~3)) {
        // a chosen wave count may not shrink the tile below most cells' size
        // (each overfull tile takes a second pass over its rows)
        const int tmx = c->tile_max;
        delete c;
        set_error("PP with %d waves: LDS tile of %d particles for a mean tile of %.0f", req[0], tmx, mean_tile);
        return MDF_EINVAL;
    }

    // PM blocks: the largest shape (up to 4 reference cells per axis) whose
    // brick fits the LDS budget; a brick spans ceil(sb * w / h) + 6 nodes per
    // axis for sb sub-cells of width w (the base-index range of the block and
    // the 6-node stencil; a margin node below for rounding at the block face
    // measured ~6% slower assignment, from the larger brick to clear and flush)
    {
        // shapes in sub-cells per axis: whole cells (sb = 2 cb) for both
        // kernels, then for the gather half a cell along x (and y)
        // (8 and 9: 3 x 3 x 2 and 3 x 2 x 2 cells, for the gather where a 3^3-cell
        // brick just misses the LDS: at N = 1e6 on 22 cells a side its
        // 25 x 25 x 24 FP32 brick takes 182 KB)
        // (10: 4 x 4 x 4 sub-cells without the LDS brick: the gather reads the
        // field mesh through the caches)
        static const int shapes[10][3] = {{8, 8, 8}, {6, 6, 6}, {4, 4, 4}, {2, 2, 2}, {1, 2, 2},
                                          {1, 1, 2}, {1, 1, 1}, {6, 6, 4}, {6, 4, 4}, {4, 4, 4}};
        auto extent = [&](const int *sb, int *B) {
            size_t nn = 1;
            for (int d = 0; d < 3; ++d) {
                const int sbd = sb[d] < g.nsub[d] ? sb[d] : g.nsub[d];
                B[d] = (int)ceil(sbd * g.w[d] / g.h[d]) + 6;
                // odd x and y extents: the row and z-plane strides of the brick
                // are then odd multiples of the element size, so the four lanes
                // that work on the rows and planes of one particle spread over
                // the LDS banks
                if (d < 2) B[d] |= 1;
                nn *= (size_t)B[d];
            }
            return nn;
        };
        auto nblocks = [&](const int *sb) {
            return ((g.nsub[0] + sb[0] - 1) / sb[0]) * ((g.nsub[1] + sb[1] - 1) / sb[1]) *
                   ((g.nsub[2] + sb[2] - 1) / sb[2]);
        };
        {
            // sized by node count (8192 nodes) in both precisions: a larger
            // block would leave too few workgroups to fill the chip; whole
            // cells only (2^3 sub-cells at the least)
            int k = 0;
            int B[3];
            while (k < 3 && extent(shapes[k], B) * 8 > 64 * 1024) ++k;
            if (req[1]) {
                k = req[1] - 1;
                if (extent(shapes[k], B) * 8 > 128 * 1024) {
                    delete c;
                    set_error("assignment block shape %d: its brick exceeds the 128 KB of LDS", req[1]);
                    return MDF_EINVAL;
                }
            }
            const size_t nn = extent(shapes[k], B);
            for (int d = 0; d < 3; ++d) {
                c->brick.sb_assign[d] = shapes[k][d];
                c->brick.B[d] = B[d];
            }
            c->brick.lds_assign = nn * 8 <= 128 * 1024 ? 1 : 0;
            c->shape_assign = k + 1;
        }
        {
            // 156 KB: with the gather's < 1 KB of static LDS within the 160 KB
            // of a CU (at 144 KB the FP64 brick of 2^3 cells at N = 1e6, rc 7,
            // 152 KB, fell back to 1-cell blocks: 0.45 ms against 0.27 ms).
            // When no shape fits, 2^3 sub-cells read the global field meshes.
            // The FP32 brick's planes are padded (gather_plane_f32); the LDS
            // left beside it holds the particles' visit order, 2 B each (up to
            // 4096, four a thread; a block with more keeps the sorted order).
            auto brick_bytes = [&](const int *sb, int *B) {
                extent(sb, B);
                return gather_brick_bytes(B, mixed);
            };
            // the largest block (in sub-cells) whose brick fits
            static const int by_size[9] = {0, 1, 7, 8, 2, 3, 4, 5, 6};
            int o = 0;
            int B[3];
            while (o < 9 && brick_bytes(shapes[by_size[o]], B) > 156 * 1024) ++o;
            bool fits = o < 9;
            int k = fits ? by_size[o] : 3;
            if (req[2] == 10) {
                k = 9;
                fits = false;
            } else if (req[2]) {
                k = req[2] - 1;
                if (brick_bytes(shapes[k], B) > 156 * 1024) {
                    delete c;
                    set_error("gather block shape %d: its field brick exceeds the 156 KB of LDS", req[2]);
                    return MDF_EINVAL;
                }
                fits = true;
            }
            const size_t bb = brick_bytes(shapes[k], B);
            for (int d = 0; d < 3; ++d) {
                c->brick.sb_gather[d] = shapes[k][d];
                c->brick.Bg[d] = B[d];
            }
            c->brick.lds_gather = fits ? 1 : 0;
            c->brick.ord_cap = 0;
            if (fits && mixed && bb < 156 * 1024) {
                const size_t cap = (156 * 1024 - bb) / 2;
                c->brick.ord_cap = (int)(cap < 4 * (size_t)kGatherThreads ? cap : 4 * (size_t)kGatherThreads);
            }
            c->shape_gather = k + 1;
        }
        c->brick.nblock_assign = nblocks(c->brick.sb_assign);
        c->brick.nblock_gather = nblocks(c->brick.sb_gather);
    }

    int r = MDF_OK;
#define ALLOC(ptr, cnt) \
    if (r == MDF_OK) r = dmalloc(&(ptr), (cnt))
    ALLOC(c->cell_count, (size_t)c->ncell);
    ALLOC(c->cell_start, (size_t)c->ncell + 1);
    ALLOC(c->scan_tmp, (size_t)c->ncell / 1024 + 2);
    ALLOC(c->part_cell, (size_t)p->n_max);
    ALLOC(c->part_slot, (size_t)p->n_max);
    ALLOC(c->perm, (size_t)p->n_max);
    ALLOC(c->spos, (size_t)p->n_max);
    ALLOC(c->acc_pp_s, 3 * (size_t)p->n_max);
    ALLOC(c->rho, c->nmesh);
    ALLOC(c->green, c->nspec);
    if (mixed) {
        ALLOC(c->chebf, pair_tab.size());
        ALLOC(c->stab, s_tab.size());
        ALLOC(c->ckf, c->nspec);
        ALLOC(c->ekf, 3 * c->nspec);
        ALLOC(c->emeshf, 3 * c->nmesh);
        ALLOC(c->greenf, c->nspec);
        for (int d = 0; d < 3; ++d) ALLOC(c->twf[d], (size_t)g.M[d] / 2);
    } else {
        ALLOC(c->cheb, pair_tab.size());
        ALLOC(c->ck, c->nspec);
        ALLOC(c->ek, 3 * c->nspec);
        ALLOC(c->emesh, 3 * c->nmesh);
        for (int d = 0; d < 3; ++d) ALLOC(c->tw[d], (size_t)g.M[d] / 2);
    }
    ALLOC(c->part_pp, (size_t)c->nblock);
    ALLOC(c->part_pm, (size_t)g.M[1] * (g.Mxh + 3));
    ALLOC(c->part_red, (size_t)kMaxPartials);
    ALLOC(c->scal, (size_t)kScalCount);
    ALLOC(c->stats, 4);
#undef ALLOC
    if (r != MDF_OK) {
        free_ctx(c);
        return r;
    }
    hipError_t up;
    if (mixed) {
        std::vector<float2> tf(pair_tab.size());
        for (size_t i = 0; i < tf.size(); ++i) tf[i] = make_float2((float)pair_tab[i].x, (float)pair_tab[i].y);
        up = hipMemcpy(c->chebf, tf.data(), tf.size() * sizeof(float2), hipMemcpyHostToDevice);
        if (up == hipSuccess) {
            std::vector<float2> ts(s_tab.size());
            for (size_t i = 0; i < ts.size(); ++i) ts[i] = make_float2((float)s_tab[i].x, (float)s_tab[i].y);
            up = hipMemcpy(c->stab, ts.data(), ts.size() * sizeof(float2), hipMemcpyHostToDevice);
        }
    } else {
        up = hipMemcpy(c->cheb, pair_tab.data(), pair_tab.size() * sizeof(double2), hipMemcpyHostToDevice);
    }
    if (up != hipSuccess) {
        set_error("pair-table upload failed");
        free_ctx(c);
        return MDF_EHIP;
    }
    // twiddles exp(-2 pi i t / M), evaluated in long double
    for (int d = 0; d < 3; ++d) {
        const int m = g.M[d];
        std::vector<double2> tw(m / 2);
        std::vector<float2> twf(m / 2);
        for (int t = 0; t < m / 2; ++t) {
            const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)m;
            tw[t] = make_double2((double)cosl(ang), (double)-sinl(ang));
            twf[t] = make_float2((float)cosl(ang), (float)-sinl(ang));
        }
        const hipError_t e = mixed ? hipMemcpy(c->twf[d], twf.data(), twf.size() * sizeof(float2), hipMemcpyHostToDevice)
                                   : hipMemcpy(c->tw[d], tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error("twiddle upload failed");
            free_ctx(c);
            return MDF_EHIP;
        }
    }
    if ((r = fill_zero_i(c->cell_count, (size_t)c->ncell, nullptr)) != MDF_OK ||
        (r = fill_zero_i(c->stats, 4, nullptr)) != MDF_OK || (r = fill_zero(c->scal, kScalCount, nullptr)) != MDF_OK ||
        (r = fill_zero(c->green, c->nspec, nullptr)) != MDF_OK) {
        free_ctx(c);
        return r;
    }
    if ((r = vlist_alloc(c, list_skin_of(c, skin, list_cap))) != MDF_OK) {
        free_ctx(c);
        return r;
    }
    c->list_skin_req = list_cap;
    c->req_waves = req[0];
    c->req_assign = req[1];
    c->req_gather = req[2];
    c->skin_req = skin;
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("device synchronisation failed after create");
        free_ctx(c);
        return MDF_EHIP;
    }
    *out = c;
    return MDF_OK;
}

}  // namespace mdf

extern "C" {

int mdf_p3m_create(const mdf_p3m_params *p, mdf_p3m_ctx **out) {
    static const int automatic[3] = {0, 0, 0};
    return create_ctx(p, 0.0, automatic, 0.0, out);
}

}  // extern "C"

namespace mdf {
// The lists' skin of a context whose cells were sized for rc + skin: the
// largest the cells allow (every pair within the smallest cell width, at
// least rc + skin, is inside the stencil), or at most cap when cap > 0.
static double list_skin_of(const mdf_p3m_ctx *c, double skin, double cap) {
    if (!(skin > 0.0)) return 0.0;
    double wmin = 2.0 * c->g.w[0];
    for (int d = 1; d < 3; ++d) wmin = std::min(wmin, 2.0 * c->g.w[d]);
    const double s = wmin - c->g.rc;
    return cap > 0.0 && cap < s ? cap : s;
}

// Rebuild a context with another skin or launch choice: a new context is
// created, the Green function and the attached network carry over, then the
// two swap contents so that the caller's handle stays valid.  On failure the
// context is unchanged.
static int rebuild_ctx(mdf_p3m_ctx *ctx, double skin, const int req[3], const char *who) {
    mdf_p3m_ctx *nc = nullptr;
    MDF_TRY(create_ctx(&ctx->prm, skin, req, ctx->list_skin_req, &nc));
    if (ctx->have_green) {
        hipError_t e = hipMemcpy(nc->green, ctx->green, ctx->nspec * sizeof(double), hipMemcpyDeviceToDevice);
        if (e == hipSuccess && ctx->greenf)
            e = hipMemcpy(nc->greenf, ctx->greenf, ctx->nspec * sizeof(float), hipMemcpyDeviceToDevice);
        if (e != hipSuccess) {
            free_ctx(nc);
            set_error("%s: Green function copy failed (%s)", who, hipGetErrorString(e));
            return MDF_EHIP;
        }
        nc->have_green = true;
    }
    nc->mlp = ctx->mlp;
    nc->gen = ctx->gen + 1;
    (void)hipDeviceSynchronize();
    mdf_p3m_ctx tmp;
    std::memcpy((void *)&tmp, (void *)ctx, sizeof tmp);
    std::memcpy((void *)ctx, (void *)nc, sizeof tmp);
    std::memcpy((void *)nc, (void *)&tmp, sizeof tmp);
    free_ctx(nc);
    return MDF_OK;
}
}  // namespace mdf

extern "C" {

int mdf_p3m_set_skin(mdf_p3m_ctx *ctx, double skin) {
    if (!ctx) {
        set_error("mdf_p3m_set_skin: null context");
        return MDF_EINVAL;
    }
    if (skin == ctx->skin_req) return MDF_OK;
    const int req[3] = {ctx->req_waves, ctx->req_assign, ctx->req_gather};
    return rebuild_ctx(ctx, skin, req, "mdf_p3m_set_skin");
}

int mdf_p3m_set_list_skin(mdf_p3m_ctx *ctx, double list_skin, double *effective) {
    if (!ctx || !(list_skin >= 0.0) || !(list_skin < 1e300)) {
        set_error("mdf_p3m_set_list_skin: need a context and a finite list skin >= 0");
        return MDF_EINVAL;
    }
    ctx->list_skin_req = list_skin;
    const double s = list_skin_of(ctx, ctx->skin_req, list_skin);
    if (s != ctx->vl_skin) {
        MDF_HIP(hipDeviceSynchronize());  // the lists of earlier calls are released
        MDF_TRY(vlist_alloc(ctx, s));
        ctx->gen += 1;
    }
    if (effective) *effective = ctx->vl_skin;
    return MDF_OK;
}

int mdf_p3m_set_launch(mdf_p3m_ctx *ctx, int pp_waves, int assign_shape, int gather_shape) {
    if (!ctx) {
        set_error("mdf_p3m_set_launch: null context");
        return MDF_EINVAL;
    }
    if (pp_waves == ctx->req_waves && assign_shape == ctx->req_assign && gather_shape == ctx->req_gather)
        return MDF_OK;
    const int req[3] = {pp_waves, assign_shape, gather_shape};
    return rebuild_ctx(ctx, ctx->skin_req, req, "mdf_p3m_set_launch");
}

int mdf_p3m_launch_config(mdf_p3m_ctx *ctx, int *out6) {
    if (!ctx || !out6) {
        set_error("mdf_p3m_launch_config: null argument");
        return MDF_EINVAL;
    }
    out6[0] = ctx->pp_waves;
    out6[1] = ctx->shape_assign;
    out6[2] = ctx->shape_gather;
    out6[3] = ctx->brick.nblock_assign;
    out6[4] = ctx->brick.nblock_gather;
    out6[5] = ctx->tile_max;
    return MDF_OK;
}

int mdf_p3m_list_stats(mdf_p3m_ctx *ctx, int *out4) {
    if (!ctx || !out4) {
        set_error("mdf_p3m_list_stats: null argument");
        return MDF_EINVAL;
    }
    int h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ctx->vl_flags) MDF_HIP(hipMemcpy(h, ctx->vl_flags, sizeof h, hipMemcpyDeviceToHost));
    out4[0] = h[4];  // builds
    out4[1] = h[5];  // steps on the lists
    out4[2] = ctx->vl_cap;
    out4[3] = ctx->g.nblk[0];
    return MDF_OK;
}

}  // extern "C"

extern "C" {

int mdf_p3m_destroy(mdf_p3m_ctx *ctx) {
    if (!ctx) return MDF_OK;
    (void)hipDeviceSynchronize();
    free_ctx(ctx);
    return MDF_OK;
}

int mdf_p3m_set_green(mdf_p3m_ctx *ctx, const double *G_k, void *stream) {
    if (!ctx || !G_k) {
        set_error("mdf_p3m_set_green: null argument");
        return MDF_EINVAL;
    }
    return set_green(ctx, G_k, (hipStream_t)stream);
}

int mdf_p3m_build_green(mdf_p3m_ctx *ctx, int p, int mx_max, int my_max, int mz_max, double *G_k_out,
                        double *A_pm_out, void *stream) {
    if (!ctx) {
        set_error("mdf_p3m_build_green: null context");
        return MDF_EINVAL;
    }
    return build_green(ctx, p, mx_max, my_max, mz_max, G_k_out, A_pm_out, (hipStream_t)stream);
}

int mdf_p3m_pm_error_sum(mdf_p3m_ctx *ctx, int p, int mx_max, int my_max, int mz_max, double *Q_sum_out,
                         void *stream) {
    if (!ctx) {
        set_error("mdf_p3m_pm_error_sum: null context");
        return MDF_EINVAL;
    }
    return pm_error_sum(ctx, p, mx_max, my_max, mz_max, Q_sum_out, (hipStream_t)stream);
}

int mdf_p3m_force(mdf_p3m_ctx *ctx, const double *pos, const double *Z, int n, double af, double uf, double *acc,
                  double *energies, double *acc_pp, double *acc_pm, double *e_x, double *e_y, double *e_z,
                  double *rho, void *stream) {
    if (!ctx || (!pos && n > 0)) {
        set_error("mdf_p3m_force: null argument");
        return MDF_EINVAL;
    }
    MDF_TRY(check_n(ctx, n));
    MDF_TRY(need_green(ctx));
    hipStream_t s = (hipStream_t)stream;
    prof_mark(ctx, kStageBin, s);
    ctx->vl_force_build = 1;  // positions from the caller: build the lists anew
    MDF_TRY(bin_step(ctx, pos, Z, n, s));
    prof_mark(ctx, kStagePP, s);
    MDF_TRY(pp_forces(ctx, n, s));
    if (ctx->mlp) MDF_TRY(mlp_correct(ctx->mlp, ctx, n, s));
    prof_mark(ctx, kStageAssign, s);
    MDF_TRY(pm_assign(ctx, n, rho, s));
    prof_mark(ctx, kStageFFT, s);
    MDF_TRY(pm_solve(ctx, s));
    GatherOut o{};
    o.acc = acc;
    o.af = af;
    o.include_pp = 1;
    o.acc_pp = acc_pp;
    o.acc_pm = acc_pm;
    o.ex = e_x;
    o.ey = e_y;
    o.ez = e_z;
    prof_mark(ctx, kStageGather, s);
    MDF_TRY(pm_gather(ctx, n, o, s));
    prof_mark(ctx, kStageFinal, s);
    MDF_TRY(finalize_energy(ctx, n, uf, 1, 1, energies, s));
    prof_mark(ctx, kStageEnd, s);
    return check_overflow(ctx, ctx->mlp, s);
}

int mdf_pp_force(mdf_p3m_ctx *ctx, const double *pos, int n, double *acc_pp, double *u_pp, void *stream) {
    if (!ctx || (!pos && n > 0)) {
        set_error("mdf_pp_force: null argument");
        return MDF_EINVAL;
    }
    MDF_TRY(check_n(ctx, n));
    hipStream_t s = (hipStream_t)stream;
    ctx->vl_force_build = 1;
    MDF_TRY(bin_step(ctx, pos, nullptr, n, s));
    MDF_TRY(pp_forces(ctx, n, s));
    if (acc_pp) MDF_TRY(unpermute3(ctx, ctx->acc_pp_s, acc_pp, n, s));
    MDF_TRY(finalize_energy(ctx, n, 1.0, 1, 0, nullptr, s));
    if (u_pp) MDF_HIP(hipMemcpyAsync(u_pp, ctx->scal + kScalUpp, sizeof(double), hipMemcpyDeviceToDevice, s));
    return check_overflow(ctx, nullptr, s);
}

int mdf_pm_force(mdf_p3m_ctx *ctx, const double *pos, const double *Z, int n, double *acc_pm, double *u_pm,
                 double *e_x, double *e_y, double *e_z, double *rho, void *stream) {
    if (!ctx || (!pos && n > 0)) {
        set_error("mdf_pm_force: null argument");
        return MDF_EINVAL;
    }
    MDF_TRY(check_n(ctx, n));
    MDF_TRY(need_green(ctx));
    hipStream_t s = (hipStream_t)stream;
    MDF_TRY(bin_particles(ctx, pos, Z, n, s));
    MDF_TRY(pm_assign(ctx, n, rho, s));
    MDF_TRY(pm_solve(ctx, s));
    GatherOut o{};
    o.acc = acc_pm;
    o.af = 1.0;
    o.include_pp = 0;
    o.ex = e_x;
    o.ey = e_y;
    o.ez = e_z;
    MDF_TRY(pm_gather(ctx, n, o, s));
    MDF_TRY(finalize_energy(ctx, n, 1.0, 0, 1, nullptr, s));
    if (u_pm) MDF_HIP(hipMemcpyAsync(u_pm, ctx->scal + kScalUpm, sizeof(double), hipMemcpyDeviceToDevice, s));
    return MDF_OK;
}

int mdf_verlet_kick_drift(mdf_p3m_ctx *ctx, double *pos, double *vel, const double *acc, int n, double dt, int pbc,
                          const double *lmin, const double *lmax, void *stream) {
    if (!ctx || (n > 0 && (!pos || !vel || !acc))) {
        set_error("mdf_verlet_kick_drift: null argument");
        return MDF_EINVAL;
    }
    MDF_TRY(check_n(ctx, n));
    return kick_drift(ctx, pos, vel, acc, n, dt, pbc, lmin, lmax, nullptr, (hipStream_t)stream);
}

int mdf_verlet_kick(mdf_p3m_ctx *ctx, double *vel, const double *acc, int n, double dt, double *k_sum,
                    void *stream) {
    if (!ctx || (n > 0 && (!vel || !acc))) {
        set_error("mdf_verlet_kick: null argument");
        return MDF_EINVAL;
    }
    MDF_TRY(check_n(ctx, n));
    return kick(ctx, vel, acc, n, dt, k_sum, (hipStream_t)stream);
}

int mdf_thermostat_scale(mdf_p3m_ctx *ctx, double *vel, int n, const double *k_sum, double T_desired, int it,
                         double *t_out, void *stream) {
    if (!ctx || !k_sum || (n > 0 && !vel)) {
        set_error("mdf_thermostat_scale: null argument");
        return MDF_EINVAL;
    }
    MDF_TRY(check_n(ctx, n));
    return thermostat_scale(ctx, vel, n, k_sum, T_desired, it, t_out, (hipStream_t)stream);
}

int mdf_density_modes(mdf_p3m_ctx *ctx, const double *pos, int n, const double *q, int nq, double *out,
                      void *stream) {
    if (!ctx || (nq > 0 && (!q || !out)) || (n > 0 && !pos)) {
        set_error("mdf_density_modes: null argument");
        return MDF_EINVAL;
    }
    MDF_TRY(check_n(ctx, n));
    return density_modes(ctx, pos, n, q, nq, out, (hipStream_t)stream);
}

int mdf_md_run(mdf_p3m_ctx *ctx, double *pos, double *vel, double *acc, const double *Z, int n, int nsteps, double dt,
               double af, double uf, double kf, const double *q, int nq, double q_step, double *energies,
               double *nqt, void *stream) {
    if (!ctx || (n > 0 && (!pos || !vel || !acc)) || nsteps < 0 || (nq > 0 && ((!q && !(q_step > 0.0)) || !nqt))) {
        set_error("mdf_md_run: bad argument");
        return MDF_EINVAL;
    }
    MDF_TRY(check_n(ctx, n));
    MDF_TRY(need_green(ctx));
    hipStream_t s = (hipStream_t)stream;
    ctx->vl_force_build = 1;  // the caller may have moved the particles since the last call
    for (int it = 0; it < nsteps; ++it) {
        MDF_TRY(md_step(ctx, pos, vel, acc, Z, n, dt, af, uf, kf, energies ? energies + 4 * (size_t)it : nullptr, s));
        if (nq > 0) {
            prof_mark(ctx, kStageModes, s);
            double *dst = nqt + (size_t)it * nq * 6;
            if (q_step > 0.0) MDF_TRY(density_modes_grid(ctx, pos, n, q_step, nq, dst, s));
            else MDF_TRY(density_modes(ctx, pos, n, q, nq, dst, s));
        }
        prof_mark(ctx, kStageEnd, s);
    }
    return check_overflow(ctx, ctx->mlp, s);
}

int mdf_p3m_profile(mdf_p3m_ctx *ctx, int mode, double *out, int nout) {
    if (!ctx) {
        set_error("mdf_p3m_profile: null context");
        return MDF_EINVAL;
    }
    if (mode == 0) {
        prof_free(ctx);
        return MDF_OK;
    }
    if (mode == 1 || mode == 3) {
        prof_free(ctx);
        Prof *p = new (std::nothrow) Prof();
        if (!p) {
            set_error("out of host memory");
            return MDF_ENOMEM;
        }
        const size_t cap = nout > 0 ? (size_t)nout : 65536;
        p->ev.resize(cap);
        p->stage.resize(cap);
        for (size_t i = 0; i < cap; ++i) {
            if (hipEventCreate(&p->ev[i]) != hipSuccess) {
                p->ev.resize(i);
                p->stage.resize(i);
                ctx->prof = p;
                prof_free(ctx);
                set_error("hipEventCreate failed");
                return MDF_EHIP;
            }
        }
        p->pp_only = mode == 3;
        ctx->prof = p;
        return MDF_OK;
    }
    if (mode == 2) {
        Prof *p = ctx->prof;
        if (!p || !out || nout < 2 * kStageCount) {
            set_error("mdf_p3m_profile: profiling off or output too small (need %d doubles)", 2 * kStageCount);
            return MDF_EINVAL;
        }
        for (int i = 0; i < 2 * kStageCount; ++i) out[i] = 0.0;
        if (p->used) MDF_HIP(hipEventSynchronize(p->ev[p->used - 1]));
        for (size_t i = 0; i + 1 < p->used; ++i) {
            const int st = p->stage[i];
            if (st < 0 || st >= kStageCount) continue;
            float ms = 0.f;
            MDF_HIP(hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]));
            out[st] += ms;
            out[kStageCount + st] += 1.0;
        }
        p->used = 0;
        if (p->overflow) {
            p->overflow = false;
            set_error("profile event pool overflowed; totals are partial");
            return MDF_ESTATE;
        }
        return MDF_OK;
    }
    set_error("mdf_p3m_profile: mode must be 0, 1, 2 or 3");
    return MDF_EINVAL;
}

int mdf_p3m_config(mdf_p3m_ctx *ctx, int *out4) {
    if (!ctx || !out4) {
        set_error("mdf_p3m_config: null argument");
        return MDF_EINVAL;
    }
    out4[0] = ctx->pp_waves;
    out4[1] = ctx->tile_max;
    out4[2] = ctx->npiece;
    out4[3] = (int)ctx->pp_lds;
    return MDF_OK;
}

int mdf_p3m_stats(mdf_p3m_ctx *ctx, int *out2) {
    if (!ctx || !out2) {
        set_error("mdf_p3m_stats: null argument");
        return MDF_EINVAL;
    }
    int h[4];
    MDF_HIP(hipMemcpy(h, ctx->stats, sizeof h, hipMemcpyDeviceToHost));
    out2[0] = h[0];
    out2[1] = h[1];
    if (h[2]) {
        set_error("a PP tile row exceeded the LDS tile capacity; forces of that step are incomplete");
        return MDF_ESTATE;
    }
    return MDF_OK;
}

}  // extern "C"

namespace mdf {

__global__ void k_unpermute3(const double *__restrict__ src, const int *__restrict__ perm, double *__restrict__ dst,
                             int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = perm[s];
    dst[3 * (size_t)i + 0] = src[3 * (size_t)s + 0];
    dst[3 * (size_t)i + 1] = src[3 * (size_t)s + 1];
    dst[3 * (size_t)i + 2] = src[3 * (size_t)s + 2];
}

int unpermute3(mdf_p3m_ctx *c, const double *src_sorted, double *dst, int n, hipStream_t s) {
    if (n <= 0) return MDF_OK;
    hipLaunchKernelGGL(k_unpermute3, dim3((n + 255) / 256), dim3(256), 0, s, src_sorted, c->perm, dst, n);
    MDF_LAUNCHED();
    return MDF_OK;
}

}  // namespace mdf
