// Replica ensembles: many independent P3M systems stepped by one graph launch.
//
// The S(q, w) averaging of s_of_q_w_avg.py needs many independent seeded
// runs of the NVE loop of yukawa_MD_p3m_ord6.py:195-219.  Here each replica
// keeps its own context (its own box, mesh, alpha, rc and scratch), and the
// steps of all replicas are captured once into a hipGraph:
//
//   caller stream --fork--> W worker streams (replica r on worker r % W)
//        each worker: K x [kick-drift, bin, PP, assign, FFT solve, gather +
//                          kick, FP64 reductions, n(q) on the q grid]
//   <--join-- flush kernel (ring -> outputs at the launch counter), counter++
//
// One graph launch advances every replica by K steps.  Two graph shapes:
//   segmented (default): every stage of the step is ONE launch for all
//     replicas; block b belongs to the replica whose block range holds it
//     (a prefix table), and each replica's kernel arguments come from a
//     device table built once before capture.  ~18 launches per step
//     whatever the number of replicas.
//   per-replica streams (mode 1, or when a replica cannot be segmented):
//     each replica's own launches on one of W worker streams, forked and
//     joined inside the graph.  Per-step energies and n(q, t) go to ring buffers of K
// steps inside the graph, and the flush kernel copies them to the caller's
// (nsteps, ...) arrays at offset launch * K read from a device counter, so
// the same executable graph is replayed without changing any argument.
// Steps beyond the last whole graph run eagerly on the caller stream.
#include <algorithm>
#include <cstring>
#include <cstdio>
#include <vector>

#include "mdf_internal.h"

struct mdf_batch {
    int nrep = 0;
    std::vector<mdf_p3m_ctx *> ctx;
    std::vector<hipStream_t> workers;
    hipStream_t cap = nullptr;  // capture origin (the legacy NULL stream cannot be captured)
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
    // captured graph and the arguments it was captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int K = 0, nq = 0;
    double dt = 0, af = 0, uf = 0, kf = 0;
    std::vector<double *> cap_pos, cap_vel, cap_acc, cap_e, cap_q;
    std::vector<int> cap_n;
    std::vector<int> cap_gen;  // the replicas' rebuild generations at capture
    // device ring buffers and flush table
    double *ring_e = nullptr;  // nrep * K * 4
    double *ring_q = nullptr;  // nrep * K * nq * 6
    size_t ring_e_cap = 0, ring_q_cap = 0;
    struct Flush {
        double *e_out;
        double *q_out;
    };
    Flush *flush = nullptr;    // nrep entries (device)
    int *counter = nullptr;    // graph launches since the run began (device)
    int mode = 0;              // 0: segmented when possible, 1: per-replica streams
    int used_mode = -1;        // shape of the captured graph
    // segmented launch tables
    std::vector<mdf::SegStage> stages;
    unsigned char *seg_mem = nullptr;
    struct StageRun {
        int stage;
        const void *tab;
        mdf::SegGrid S;
        int total;
        size_t lds;
    };
    std::vector<StageRun> runs;
};

namespace mdf {

int md_step(mdf_p3m_ctx *c, double *pos, double *vel, double *acc, const double *Z, int n, double dt, double af,
            double uf, double kf, double *rec, hipStream_t s);

__global__ void k_batch_flush(const mdf_batch::Flush *__restrict__ tab, const double *__restrict__ ring_e,
                              const double *__restrict__ ring_q, const int *__restrict__ counter, int K, int nq) {
    const int r = blockIdx.x;
    const int launch = *counter;
    const mdf_batch::Flush f = tab[r];
    const size_t ne = (size_t)K * 4, nqv = (size_t)K * nq * 6;
    if (f.e_out) {
        const double *src = ring_e + (size_t)r * ne;
        double *dst = f.e_out + (size_t)launch * ne;
        for (size_t i = threadIdx.x; i < ne; i += blockDim.x) dst[i] = src[i];
    }
    if (f.q_out && nq > 0) {
        const double *src = ring_q + (size_t)r * nqv;
        double *dst = f.q_out + (size_t)launch * nqv;
        for (size_t i = threadIdx.x; i < nqv; i += blockDim.x) dst[i] = src[i];
    }
}

__global__ void k_batch_count(int *counter, int reset) {
    if (threadIdx.x == 0) *counter = reset ? 0 : *counter + 1;
}

static void free_graph(mdf_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static void free_batch(mdf_batch *b) {
    if (!b) return;
    free_graph(b);
    for (mdf_p3m_ctx *c : b->ctx)
        if (c) mdf_p3m_destroy(c);
    for (hipStream_t w : b->workers) (void)hipStreamDestroy(w);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    for (hipEvent_t e : b->ev_join) (void)hipEventDestroy(e);
    if (b->ring_e) (void)hipFree(b->ring_e);
    if (b->ring_q) (void)hipFree(b->ring_q);
    if (b->flush) (void)hipFree(b->flush);
    if (b->seg_mem) (void)hipFree(b->seg_mem);
    if (b->counter) (void)hipFree(b->counter);
    delete b;
}

static double q_step_of(const mdf_p3m_ctx *c) { return 2.0 * M_PI / c->g.L[0]; }

// Enqueue K steps of replica r on stream w, records into the ring.
static int enqueue_replica(mdf_batch *b, int r, int K, hipStream_t w) {
    mdf_p3m_ctx *c = b->ctx[r];
    for (int k = 0; k < K; ++k) {
        double *rec = b->ring_e + ((size_t)r * K + k) * 4;
        MDF_TRY(md_step(c, b->cap_pos[r], b->cap_vel[r], b->cap_acc[r], nullptr, b->cap_n[r], b->dt, b->af, b->uf,
                        b->kf, rec, w));
        if (b->nq > 0) {
            double *dst = b->ring_q + (((size_t)r * K + k) * b->nq) * 6;
            MDF_TRY(density_modes_grid(c, b->cap_pos[r], b->cap_n[r], q_step_of(c), b->nq, dst, w));
        }
    }
    return MDF_OK;
}

// Builds every stage's argument table for all replicas and uploads them
// (synchronously, before capture).  Replicas whose dynamic LDS differ by more
// than 25% go to separate launches of the stage (each launch reserves its
// largest replica's LDS for all its workgroups, so one replica with a large
// tile would otherwise lower the occupancy of all).  Fails if a replica
// cannot be segmented.
static int seg_prepare(mdf_batch *b) {
    const int R = b->nrep;
    const int prec = b->ctx[0]->prm.precision;
    for (int r = 0; r < R; ++r) {
        if (b->ctx[r]->prm.precision != prec) {
            set_error("replica batch: mixed precisions");
            return MDF_EINVAL;
        }
        if (b->ctx[r]->mlp) {
            set_error("replica batch: a replica has a learned correction attached");
            return MDF_EINVAL;
        }
    }
    seg_step_stages(prec, b->stages);
    if (b->nq == 0) b->stages.resize(b->stages.size() - 2);  // no n(q) stages
    const size_t nst = b->stages.size();
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    struct Group {
        size_t stage;
        std::vector<int> reps;
        size_t lds = 0;
        size_t off = 0;
        int total = 0;
        int variant = 0;
    };
    std::vector<Group> groups;
    std::vector<std::vector<unsigned char>> args(nst);
    std::vector<std::vector<int4>> dims(nst, std::vector<int4>(R));
    std::vector<std::vector<size_t>> lds(nst, std::vector<size_t>(R, 0));
    for (size_t i = 0; i < nst; ++i) {
        const SegStage &st = b->stages[i];
        args[i].assign(st.arg_size * R, 0);
        for (int r = 0; r < R; ++r) {
            StepIO io;
            io.pos = b->cap_pos[r];
            io.vel = b->cap_vel[r];
            io.acc = b->cap_acc[r];
            io.n = b->cap_n[r];
            io.dt = b->dt;
            io.af = b->af;
            io.uf = b->uf;
            io.kf = b->kf;
            io.rec0 = b->ring_e + (size_t)r * b->K * 4;
            io.rec_stride = 4;
            io.modes0 = b->ring_q + (size_t)r * b->K * (b->nq > 0 ? b->nq : 1) * 6;
            io.modes_stride = b->nq * 6;
            io.nq = b->nq;
            const int rc = st.build(b->ctx[r], io, args[i].data() + st.arg_size * r, dims[i][r], lds[i][r]);
            if (rc != MDF_OK) return rc;
        }
        // groups of one kernel variant (dims.w) and similar LDS, in order of
        // variant and LDS
        std::vector<int> order(R);
        for (int r = 0; r < R; ++r) order[r] = r;
        std::sort(order.begin(), order.end(), [&](int a, int c) {
            return dims[i][a].w != dims[i][c].w ? dims[i][a].w < dims[i][c].w : lds[i][a] < lds[i][c];
        });
        Group g;
        g.stage = i;
        size_t gmin = 0;
        for (int r : order) {
            if (!g.reps.empty() && (lds[i][r] > gmin + gmin / 4 + 1024 || dims[i][r].w != g.variant)) {
                groups.push_back(g);
                g = Group();
                g.stage = i;
            }
            if (g.reps.empty()) {
                gmin = lds[i][r];
                g.variant = dims[i][r].w;
            }
            g.reps.push_back(r);
            if (lds[i][r] > g.lds) g.lds = lds[i][r];
        }
        groups.push_back(g);
    }
    size_t bytes = 0;
    for (Group &g : groups) {
        const size_t n = g.reps.size();
        g.off = bytes;
        bytes += align(b->stages[g.stage].arg_size * n) + align(sizeof(int) * (n + 1)) + align(sizeof(int4) * n);
    }
    std::vector<unsigned char> host(bytes, 0);
    for (Group &g : groups) {
        const SegStage &st = b->stages[g.stage];
        const size_t n = g.reps.size();
        unsigned char *base = host.data() + g.off;
        int *first = (int *)(base + align(st.arg_size * n));
        int4 *dd = (int4 *)(base + align(st.arg_size * n) + align(sizeof(int) * (n + 1)));
        int acc = 0;
        for (size_t q = 0; q < n; ++q) {
            const int r = g.reps[q];
            std::memcpy(base + st.arg_size * q, args[g.stage].data() + st.arg_size * r, st.arg_size);
            const int4 d = dims[g.stage][r];
            dd[q] = d;
            first[q] = acc;
            acc += d.x * d.y * d.z;
        }
        first[n] = acc;
        g.total = acc;
    }
    if (b->seg_mem) (void)hipFree(b->seg_mem);
    b->seg_mem = nullptr;
    MDF_HIP(hipMalloc((void **)&b->seg_mem, bytes));
    MDF_HIP(hipMemcpy(b->seg_mem, host.data(), bytes, hipMemcpyHostToDevice));
    b->runs.clear();
    for (const Group &g : groups) {
        const SegStage &st = b->stages[g.stage];
        const size_t n = g.reps.size();
        unsigned char *base = b->seg_mem + g.off;
        mdf_batch::StageRun u;
        u.stage = (int)g.stage;
        u.tab = base;
        u.S.first = (const int *)(base + align(st.arg_size * n));
        u.S.dims = (const int4 *)(base + align(st.arg_size * n) + align(sizeof(int) * (n + 1)));
        u.S.nrep = (int)n;
        u.S.variant = g.variant;
        u.total = g.total;
        u.lds = g.lds;
        b->runs.push_back(u);
    }
    return MDF_OK;
}

static int capture_segmented(mdf_batch *b) {
    free_graph(b);
    hipStream_t s = b->cap;
    MDF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    int r = MDF_OK;
    for (int k = 0; k < b->K && r == MDF_OK; ++k)
        for (size_t i = 0; i < b->runs.size() && r == MDF_OK; ++i) {
            const mdf_batch::StageRun &u = b->runs[i];
            r = b->stages[u.stage].launch(u.tab, u.S, u.total, u.lds, k, s);
        }
    if (r == MDF_OK) {
        hipLaunchKernelGGL(k_batch_flush, dim3(b->nrep), dim3(256), 0, s, b->flush, b->ring_e, b->ring_q, b->counter,
                           b->K, b->nq);
        hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(64), 0, s, b->counter, 0);
        if (hipGetLastError() != hipSuccess) r = MDF_EHIP;
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (r != MDF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        return MDF_EHIP;
    }
    b->graph = g;
    const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(ei));
        free_graph(b);
        return MDF_EHIP;
    }
    return MDF_OK;
}

// Captures on the batch's own stream; the graph is launched on the caller's.
static int capture(mdf_batch *b) {
    free_graph(b);
    hipStream_t s = b->cap;
    const int W = (int)b->workers.size();
    MDF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    int r = MDF_OK;
    if (hipEventRecord(b->ev_fork, s) != hipSuccess) r = MDF_EHIP;
    for (int w = 0; w < W && r == MDF_OK; ++w)
        if (hipStreamWaitEvent(b->workers[w], b->ev_fork, 0) != hipSuccess) r = MDF_EHIP;
    for (int i = 0; i < b->nrep && r == MDF_OK; ++i) r = enqueue_replica(b, i, b->K, b->workers[i % W]);
    for (int w = 0; w < W && r == MDF_OK; ++w) {
        if (hipEventRecord(b->ev_join[w], b->workers[w]) != hipSuccess || hipStreamWaitEvent(s, b->ev_join[w], 0) != hipSuccess)
            r = MDF_EHIP;
    }
    if (r == MDF_OK) {
        hipLaunchKernelGGL(k_batch_flush, dim3(b->nrep), dim3(256), 0, s, b->flush, b->ring_e, b->ring_q, b->counter,
                           b->K, b->nq);
        hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(64), 0, s, b->counter, 0);
        if (hipGetLastError() != hipSuccess) r = MDF_EHIP;
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (r != MDF_OK) {
        if (g) (void)hipGraphDestroy(g);
        if (r == MDF_EHIP) set_error("batch graph capture failed");
        return r;
    }
    if (e != hipSuccess) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        return MDF_EHIP;
    }
    b->graph = g;
    const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(ei));
        free_graph(b);
        return MDF_EHIP;
    }
    return MDF_OK;
}

}  // namespace mdf

using namespace mdf;

extern "C" {

int mdf_batch_create(int nrep, const mdf_p3m_params *params, int n_streams, mdf_batch **out) {
    if (!out || !params || nrep <= 0) {
        set_error("mdf_batch_create: need nrep > 0 and non-null params/out");
        return MDF_EINVAL;
    }
    *out = nullptr;
    mdf_batch *b = new (std::nothrow) mdf_batch();
    if (!b) {
        set_error("out of host memory");
        return MDF_ENOMEM;
    }
    b->nrep = nrep;
    b->ctx.assign(nrep, nullptr);
    for (int r = 0; r < nrep; ++r) {
        const int rc = mdf_p3m_create(&params[r], &b->ctx[r]);
        if (rc != MDF_OK) {
            char msg[1024];
            snprintf(msg, sizeof msg, "replica %d: %s", r, mdf_last_error());
            free_batch(b);
            set_error("%s", msg);
            return rc;
        }
    }
    int W = n_streams > 0 ? n_streams : 8;
    if (W > nrep) W = nrep;
    for (int w = 0; w < W; ++w) {
        hipStream_t st;
        hipEvent_t ev;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            free_batch(b);
            set_error("hipStreamCreate failed");
            return MDF_EHIP;
        }
        b->workers.push_back(st);
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            free_batch(b);
            set_error("hipEventCreate failed");
            return MDF_EHIP;
        }
        b->ev_join.push_back(ev);
    }
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipMalloc((void **)&b->flush, sizeof(mdf_batch::Flush) * nrep) != hipSuccess ||
        hipMalloc((void **)&b->counter, sizeof(int)) != hipSuccess) {
        free_batch(b);
        set_error("batch set-up allocation failed");
        return MDF_ENOMEM;
    }
    *out = b;
    return MDF_OK;
}

int mdf_batch_destroy(mdf_batch *b) {
    if (!b) return MDF_OK;
    (void)hipDeviceSynchronize();
    free_batch(b);
    return MDF_OK;
}

int mdf_batch_set_mode(mdf_batch *b, int mode) {
    if (!b || mode < 0 || mode > 1) {
        set_error("mdf_batch_set_mode: mode must be 0 (segmented) or 1 (per-replica streams)");
        return MDF_EINVAL;
    }
    if (b->mode != mode) free_graph(b);
    b->mode = mode;
    return MDF_OK;
}

int mdf_batch_mode(mdf_batch *b) { return b ? b->used_mode : MDF_EINVAL; }

int mdf_batch_context(mdf_batch *b, int r, mdf_p3m_ctx **out) {
    if (!b || !out || r < 0 || r >= b->nrep) {
        set_error("mdf_batch_context: bad replica index");
        return MDF_EINVAL;
    }
    *out = b->ctx[r];
    return MDF_OK;
}

int mdf_batch_run(mdf_batch *b, double *const *pos, double *const *vel, double *const *acc, const int *n, int nsteps,
                  double dt, double af, double uf, double kf, int nq, double *const *energies, double *const *nqt,
                  int graph_steps, void *stream) {
    if (!b || !pos || !vel || !acc || !n || nsteps < 0 || nq < 0 || nq > 64 || graph_steps < 0 ||
        (nq > 0 && !nqt)) {
        set_error("mdf_batch_run: bad argument");
        return MDF_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    for (int r = 0; r < b->nrep; ++r) {
        mdf_p3m_ctx *c = b->ctx[r];
        if (n[r] < 0 || n[r] > c->prm.n_max || (n[r] > 0 && (!pos[r] || !vel[r] || !acc[r])) ||
            (nq > 0 && !nqt[r])) {
            set_error("mdf_batch_run: replica %d: bad particle count or null array", r);
            return MDF_EINVAL;
        }
        if (!c->have_green) {
            set_error("mdf_batch_run: replica %d has no Green function", r);
            return MDF_ESTATE;
        }
    }
    // replicas with pair lists build at their first step: the caller may have
    // moved the particles since the last call (inside the graph the list
    // decisions are the device's own)
    for (int r = 0; r < b->nrep; ++r) MDF_TRY(vlist_request_build(b->ctx[r], n[r], s));
    const int K = graph_steps > 0 ? (graph_steps < nsteps ? graph_steps : nsteps) : 0;
    const int nfull = K > 0 ? nsteps / K : 0;
    if (nfull > 0) {
        // (re)capture when the arguments differ from the captured ones
        bool same = b->exec && b->K == K && b->nq == nq && b->dt == dt && b->af == af && b->uf == uf &&
                    b->kf == kf;
        // (a replica rebuilt in place since, by mdf_p3m_set_skin or
        // mdf_p3m_set_launch, has new buffers: capture again)
        for (int r = 0; same && r < b->nrep; ++r)
            same = b->cap_pos[r] == pos[r] && b->cap_vel[r] == vel[r] && b->cap_acc[r] == acc[r] &&
                   b->cap_n[r] == n[r] && b->cap_e[r] == (energies ? energies[r] : nullptr) &&
                   b->cap_q[r] == (nq > 0 ? nqt[r] : nullptr) && b->cap_gen[r] == b->ctx[r]->gen;
        if (!same) {
            const size_t ne = (size_t)b->nrep * K * 4, nqn = (size_t)b->nrep * K * (nq > 0 ? nq : 1) * 6;
            if (ne > b->ring_e_cap) {
                if (b->ring_e) (void)hipFree(b->ring_e);
                b->ring_e = nullptr;
                b->ring_e_cap = 0;
                MDF_HIP(hipMalloc((void **)&b->ring_e, ne * sizeof(double)));
                b->ring_e_cap = ne;
            }
            if (nqn > b->ring_q_cap) {
                if (b->ring_q) (void)hipFree(b->ring_q);
                b->ring_q = nullptr;
                b->ring_q_cap = 0;
                MDF_HIP(hipMalloc((void **)&b->ring_q, nqn * sizeof(double)));
                b->ring_q_cap = nqn;
            }
            b->K = K;
            b->nq = nq;
            b->dt = dt;
            b->af = af;
            b->uf = uf;
            b->kf = kf;
            b->cap_pos.assign(pos, pos + b->nrep);
            b->cap_vel.assign(vel, vel + b->nrep);
            b->cap_acc.assign(acc, acc + b->nrep);
            b->cap_n.assign(n, n + b->nrep);
            b->cap_gen.resize(b->nrep);
            for (int r = 0; r < b->nrep; ++r) b->cap_gen[r] = b->ctx[r]->gen;
            b->cap_e.assign(b->nrep, nullptr);
            b->cap_q.assign(b->nrep, nullptr);
            std::vector<mdf_batch::Flush> tab(b->nrep);
            for (int r = 0; r < b->nrep; ++r) {
                b->cap_e[r] = energies ? energies[r] : nullptr;
                b->cap_q[r] = nq > 0 ? nqt[r] : nullptr;
                tab[r].e_out = b->cap_e[r];
                tab[r].q_out = b->cap_q[r];
            }
            MDF_HIP(hipMemcpy(b->flush, tab.data(), sizeof(mdf_batch::Flush) * b->nrep, hipMemcpyHostToDevice));
            b->used_mode = 1;
            if (b->mode == 0 && seg_prepare(b) == MDF_OK) {
                MDF_TRY(capture_segmented(b));
                b->used_mode = 0;
            } else {
                MDF_TRY(capture(b));
            }
        }
        hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(64), 0, s, b->counter, 1);
        MDF_LAUNCHED();
        for (int l = 0; l < nfull; ++l) MDF_HIP(hipGraphLaunch(b->exec, s));
    }
    // remaining steps, eagerly on the caller stream
    for (int it = nfull * K; it < nsteps; ++it) {
        for (int r = 0; r < b->nrep; ++r) {
            mdf_p3m_ctx *c = b->ctx[r];
            double *rec = energies && energies[r] ? energies[r] + 4 * (size_t)it : nullptr;
            MDF_TRY(md_step(c, pos[r], vel[r], acc[r], nullptr, n[r], dt, af, uf, kf, rec, s));
            if (nq > 0)
                MDF_TRY(density_modes_grid(c, pos[r], n[r], q_step_of(c), nq, nqt[r] + (size_t)it * nq * 6, s));
        }
    }
    // the graph's work ran on the replicas' own streams, joined back into s
    for (int r = 0; r < b->nrep; ++r) {
        const int rc = check_overflow(b->ctx[r], b->ctx[r]->mlp, s);
        if (rc != MDF_OK) {
            char msg[1024];
            snprintf(msg, sizeof msg, "replica %d: %s", r, mdf_last_error());
            set_error("%s", msg);
            return rc;
        }
    }
    return MDF_OK;
}

}  // extern "C"
