// Batched solves: many independent systems (one context each) advanced
// together, every kernel of an iteration launched once for all of them.
//
// Each context's own solve code enqueues its kernels through the launchers
// of bmg_kernels.hip.  With a recorder set on the thread, those launchers
// append (kernel, grid, argument struct) records instead of launching, so
// one pass of a context's solve phase yields its launch sequence.  The
// sequences of all contexts are merged in lockstep: repeatedly the kernel at
// the head of the most sequences is taken from all of them and becomes one
// launch of k_many over those systems.  Every context keeps its own order
// and the systems share no memory, so each system computes exactly what its
// own solve computes.  The merged iteration is captured in a HIP graph and
// replayed; systems that have converged are switched off by a flag the
// kernels read, so the graph never changes during a solve.  The launch
// grid of every merged launch (all its workgroups, or a fixed persistent
// grid walking them) is chosen by timing the candidates once per plan.
#include "bmg_context.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <utility>

namespace bmg {

namespace {

size_t alignUp(size_t v, size_t a) { return (v + a - 1)/a*a; }

// Sets the thread's recorder for its lifetime.
struct Recording
{
    LaunchRecorder rec;
    Recording() { launch_recorder() = &rec; }
    ~Recording() { launch_recorder() = nullptr; }
};

// One kernel launched for several systems.
struct Merged
{
    const void* kind = nullptr;
    void (*many)(const BatchLaunch&, hipStream_t) = nullptr;
    size_t shm = 0;
    size_t argSize = 0, align = 8;
    std::vector<unsigned char> args;   // argument structs, argSize apart
    std::vector<BatchItem> items;
    unsigned total = 0;                // virtual workgroups
    size_t argsOff = 0, itemsOff = 0;  // offsets in the plan's buffer
    unsigned grid = 0;                 // workgroups launched
};

}  // namespace

// A merged launch sequence with its device copy of argument structs and
// system shares.
struct BatchPlan
{
    std::vector<Merged> launches;
    unsigned char* dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (dev) (void)hipFree(dev);
    }

    // Lockstep merge of per-system sequences; the active-flag index of a
    // system is the index of its sequence.
    void build(std::vector<std::vector<LaunchRecord>>& seqs)
    {
        const size_t n = seqs.size();
        std::vector<size_t> head(n, 0);
        for (;;)
        {
            // kernels at the heads: number of sequences, longest remainder
            std::map<std::pair<const void*, size_t>,
                     std::pair<size_t, size_t>> heads;
            for (size_t s = 0; s < n; s++)
            {
                if (head[s] >= seqs[s].size()) continue;
                const LaunchRecord& r = seqs[s][head[s]];
                auto& h = heads[{r.kind, r.shm}];
                h.first++;
                h.second = std::max(h.second, seqs[s].size() - head[s]);
            }
            if (heads.empty()) break;
            auto best = heads.begin();
            for (auto it = heads.begin(); it != heads.end(); ++it)
                if (it->second > best->second) best = it;
            Merged m;
            m.kind = best->first.first;
            m.shm = best->first.second;
            for (size_t s = 0; s < n; s++)
            {
                if (head[s] >= seqs[s].size()) continue;
                const LaunchRecord& r = seqs[s][head[s]];
                if (r.kind != m.kind || r.shm != m.shm) continue;
                head[s]++;
                if (m.items.empty())
                {
                    m.many = r.many;
                    m.argSize = r.args.size();
                    m.align = r.align;
                }
                const unsigned wg = r.grid[0]*r.grid[1]*r.grid[2];
                if (wg == 0) continue;
                BatchItem it;
                it.start = m.total;
                it.gx = r.grid[0];
                it.gy = r.grid[1];
                it.n = wg;
                it.sys = int(s);
                m.items.push_back(it);
                m.args.insert(m.args.end(), r.args.begin(), r.args.end());
                // each share starts at a multiple of 8 (its XCD order)
                m.total += unsigned(alignUp(wg, 8));
            }
            if (!m.items.empty()) launches.push_back(std::move(m));
        }
        size_t bytes = 0;
        for (auto& m : launches)
        {
            bytes = alignUp(bytes, 256);
            m.argsOff = bytes;
            bytes += m.args.size();
            bytes = alignUp(bytes, 256);
            m.itemsOff = bytes;
            bytes += m.items.size()*sizeof(BatchItem);
            m.grid = m.total;
        }
        std::vector<unsigned char> host(std::max<size_t>(bytes, 1), 0);
        for (auto& m : launches)
        {
            std::memcpy(host.data() + m.argsOff, m.args.data(), m.args.size());
            std::memcpy(host.data() + m.itemsOff, m.items.data(),
                        m.items.size()*sizeof(BatchItem));
        }
        HIPCK(hipMalloc(&dev, host.size()));
        HIPCK(hipMemcpy(dev, host.data(), host.size(), hipMemcpyHostToDevice));
    }

    void launch(const Merged& m, const int* active, hipStream_t s) const
    {
        BatchLaunch b;
        b.args = dev + m.argsOff;
        b.items = reinterpret_cast<const BatchItem*>(dev + m.itemsOff);
        b.nitems = int(m.items.size());
        b.total = m.total;
        b.active = active;
        b.grid = std::max(1u, std::min(m.grid, m.total));
        b.shm = m.shm;
        m.many(b, s);
    }

    void enqueue(const int* active, hipStream_t s) const
    {
        for (const auto& m : launches) launch(m, active, s);
        HIPCK(hipGetLastError());
    }

    // Runs the sequence once, synchronously.
    void run(hipStream_t s) const
    {
        enqueue(nullptr, s);
        HIPCK(hipStreamSynchronize(s));
    }

    // Picks the grid of every launch: all of its workgroups, or a
    // persistent grid of 4, 8 or 16 workgroups per CU walking them, by the
    // fastest of a few timed runs of the whole sequence.
    void autotune(hipStream_t s, int cus)
    {
        const unsigned cand[4] = {0u, unsigned(4*cus), unsigned(8*cus),
                                  unsigned(16*cus)};
        const size_t n = launches.size();
        std::vector<hipEvent_t> ev(n + 1);
        for (auto& e : ev) HIPCK(hipEventCreate(&e));
        std::vector<float> best(n, 1e30f);
        std::vector<unsigned> pick(n, 0);
        try
        {
            for (int rep = 0; rep < 2; rep++)
                for (unsigned c : cand)
                {
                    HIPCK(hipEventRecord(ev[0], s));
                    for (size_t i = 0; i < n; i++)
                    {
                        Merged& m = launches[i];
                        m.grid = c ? std::min(c, m.total) : m.total;
                        launch(m, nullptr, s);
                        HIPCK(hipEventRecord(ev[i + 1], s));
                    }
                    HIPCK(hipGetLastError());
                    HIPCK(hipStreamSynchronize(s));
                    for (size_t i = 0; i < n; i++)
                    {
                        float ms = 0.0f;
                        HIPCK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
                        if (rep > 0 && ms < best[i])
                        {
                            best[i] = ms;
                            pick[i] = launches[i].grid;
                        }
                    }
                }
        }
        catch (...)
        {
            for (auto& e : ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        for (size_t i = 0; i < n; i++) launches[i].grid = pick[i];
    }

    void capture(const int* active, hipStream_t s)
    {
        HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try
        {
            enqueue(active, s);
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCK(hipStreamEndCapture(s, &graph));
        HIPCK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }

    size_t workgroups() const
    {
        size_t t = 0;
        for (const auto& m : launches) t += m.total;
        return t;
    }
};

}  // namespace bmg

using namespace bmg;

struct bmg_batch
{
    std::vector<bmg_ctx*> ctx;
    hipStream_t stream = nullptr;
    int* hActive = nullptr;   // pinned
    int* dActive = nullptr;
    std::unique_ptr<BatchPlan> iter;
    std::vector<long> iterGen;   // context generations the plan was built on
    std::vector<int> iterMixed;
    bool tune = true;
    int cus = 256;

    ~bmg_batch()
    {
        iter.reset();
        if (hActive) (void)hipHostFree(hActive);
        if (dActive) (void)hipFree(dActive);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename F>
std::unique_ptr<BatchPlan> recordPlan(bmg_batch* B, F&& phase)
{
    std::vector<std::vector<LaunchRecord>> seqs(B->ctx.size());
    for (size_t s = 0; s < B->ctx.size(); s++)
    {
        Recording r;
        phase(B->ctx[s], s);
        seqs[s] = std::move(r.rec.recs);
    }
    std::unique_ptr<BatchPlan> p(new BatchPlan);
    p->build(seqs);
    return p;
}

bool mixedOf(const bmg_ctx* c) { return c->opts.precision == BMG_MIXED; }

}  // namespace

namespace bmg {

void batch_create(bmg_batch** out, bmg_ctx* const* ctxs, int n)
{
    if (!out || !ctxs || n < 1) throw Error(BMG_ERR_ARG, "no contexts");
    std::unique_ptr<bmg_batch> B(new bmg_batch);
    int dev = -1;
    for (int i = 0; i < n; i++)
    {
        if (!ctxs[i]) throw Error(BMG_ERR_ARG, "null context");
        if (dev >= 0 && ctxs[i]->device != dev)
            throw Error(BMG_ERR_ARG, "contexts on different devices");
        dev = ctxs[i]->device;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i])
                throw Error(BMG_ERR_ARG, "a context appears twice");
        B->ctx.push_back(ctxs[i]);
    }
    HIPCK(hipSetDevice(dev));
    HIPCK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    HIPCK(hipHostMalloc(&B->hActive, n*sizeof(int)));
    HIPCK(hipMalloc(&B->dActive, n*sizeof(int)));
    hipDeviceProp_t prop;
    HIPCK(hipGetDeviceProperties(&prop, dev));
    B->cus = std::max(1, prop.multiProcessorCount);
    *out = B.release();
}

// MG::Solve (MG.C:132-254) of every context, as bmg_ctx::runSolve does it,
// with each phase between two host decisions recorded and run as merged
// launches.
void batch_solve(bmg_batch* B, double* const* db, double* const* dx,
                 bool constMatrix, bmg_stats* st)
{
    if (!B || !db || !dx) throw Error(BMG_ERR_ARG, "null pointer");
    const size_t n = B->ctx.size();
    for (size_t s = 0; s < n; s++)
    {
        bmg_ctx* c = B->ctx[s];
        if (!db[s] || !dx[s]) throw Error(BMG_ERR_ARG, "null field pointer");
        if (c->opts.krylov != BMG_KRYLOV_NONE)
            throw Error(BMG_ERR_ARG, "batched solves run MG iterations "
                                     "(krylov none)");
        if (c->opts.smoother != BMG_RBGS)
            throw Error(BMG_ERR_ARG, "batched solves use the rbgs smoother");
        HIPCK(hipStreamSynchronize(c->stream));
    }
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    std::exception_ptr err;
    try
    {
        HIPCK(hipEventRecord(e0, B->stream));
        // host-side preparation (first solve, or a changed matrix)
        std::vector<char> hostPrepared(n, 0);
        for (size_t s = 0; s < n; s++)
        {
            bmg_ctx* c = B->ctx[s];
            if (c->prepareNeedsHost(constMatrix))
            {
                c->prepare(constMatrix);
                HIPCK(hipStreamSynchronize(c->stream));
                hostPrepared[s] = 1;
            }
            if (mixedOf(c))
            {
                c->ensureMixed();
                HIPCK(hipStreamSynchronize(c->stream));
            }
        }
        // the merged iteration, rebuilt when a context's buffers changed
        bool rebuild = !B->iter;
        for (size_t s = 0; s < n && !rebuild; s++)
            rebuild = B->iterGen[s] != B->ctx[s]->gen ||
                      B->iterMixed[s] != int(mixedOf(B->ctx[s]));
        if (rebuild)
        {
            B->iter = recordPlan(B, [](bmg_ctx* c, size_t) {
                c->enqueueIteration(mixedOf(c));
            });
            B->iterGen.resize(n);
            B->iterMixed.resize(n);
            for (size_t s = 0; s < n; s++)
            {
                B->iterGen[s] = B->ctx[s]->gen;
                B->iterMixed[s] = int(mixedOf(B->ctx[s]));
            }
            // timing runs before the solve's fields are initialised below
            if (B->tune) B->iter->autotune(B->stream, B->cus);
            B->iter->capture(B->dActive, B->stream);
        }
        // initial residual and the sum of x
        recordPlan(B, [&](bmg_ctx* c, size_t s) {
            Level& q = c->lv[0];
            launch_ref_to_split<double>(q.L, db[s], pr(q.b), c->stream);
            launch_ref_to_split<double>(q.L, dx[s], pr(q.x), c->stream);
            if (!hostPrepared[s]) c->prepare(constMatrix);
            c->enqueueSolveStart();
        })->run(B->stream);
        // normFactors
        std::vector<double> init(n), nf(n), cur(n);
        std::vector<int> iter(n, 0), conv(n, 0);
        recordPlan(B, [&](bmg_ctx* c, size_t) {
            c->enqueueNormTerms(c->hScalar[1]/double(c->lv[0].cells()));
        })->run(B->stream);
        for (size_t s = 0; s < n; s++)
        {
            bmg_ctx* c = B->ctx[s];
            const double cells = double(c->lv[0].cells());
            nf[s] = std::max(std::sqrt(c->hScalar[2]), cells*1e-12);
            init[s] = std::sqrt(c->hScalar[0])/nf[s];
            if (!std::isfinite(init[s]))
                throw Error(BMG_ERR_NUMERIC, "initial residual is not finite");
            cur[s] = init[s];
        }
        auto check = [&](size_t s, double v) -> int
        {
            const bmg_opts& o = B->ctx[s]->opts;
            return (init[s] > 0.0 && v > o.tolerance &&
                    v/(init[s] + 1e-20) > o.rel_tol) ? 0 : 1;
        };
        for (size_t s = 0; s < n; s++) conv[s] = check(s, cur[s]);
        for (;;)
        {
            size_t nactive = 0;
            for (size_t s = 0; s < n; s++)
            {
                const bmg_opts& o = B->ctx[s]->opts;
                const bool go = (iter[s] < o.max_iter && conv[s] == 0) ||
                                iter[s] < o.min_iter;
                B->hActive[s] = go;
                nactive += go;
            }
            if (nactive == 0) break;
            HIPCK(hipMemcpyAsync(B->dActive, B->hActive, n*sizeof(int),
                                 hipMemcpyHostToDevice, B->stream));
            HIPCK(hipGraphLaunch(B->iter->exec, B->stream));
            HIPCK(hipStreamSynchronize(B->stream));
            for (size_t s = 0; s < n; s++)
            {
                if (!B->hActive[s]) continue;
                bmg_ctx* c = B->ctx[s];
                cur[s] = std::sqrt(c->hScalar[0])/nf[s];
                if (!std::isfinite(cur[s]))
                    throw Error(BMG_ERR_NUMERIC, "residual norm is not finite");
                conv[s] = check(s, cur[s]);
                iter[s]++;
            }
        }
        // boundary values of x, results back in the reference layout
        recordPlan(B, [&](bmg_ctx* c, size_t s) {
            Level& q = c->lv[0];
            c->enqueueSolveEnd();
            launch_split_to_ref<double>(q.L, cpr(q.x), dx[s], c->stream);
            launch_split_to_ref<double>(q.L, cpr(q.b), db[s], c->stream);
        })->run(B->stream);
        HIPCK(hipEventRecord(e1, B->stream));
        HIPCK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIPCK(hipEventElapsedTime(&ms, e0, e1));
        if (st)
            for (size_t s = 0; s < n; s++)
            {
                bmg_stats& o = st[s];
                std::memset(&o, 0, sizeof(o));
                o.status = BMG_OK;
                o.iters = iter[s];
                o.initial_residual = init[s];
                o.final_residual = cur[s];
                o.solve_ms = ms;
                o.learned_fallback = -1;
            }
    }
    catch (...)
    {
        err = std::current_exception();
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err) std::rethrow_exception(err);
}

void batch_info(const bmg_batch* B, int* launches, long* workgroups)
{
    if (!B) throw Error(BMG_ERR_ARG, "null batch");
    if (launches) *launches = B->iter ? int(B->iter->launches.size()) : 0;
    if (workgroups) *workgroups = B->iter ? long(B->iter->workgroups()) : 0;
}

void batch_destroy(bmg_batch* B) { delete B; }

void batch_set_autotune(bmg_batch* B, int on)
{
    if (!B) throw Error(BMG_ERR_ARG, "null batch");
    B->tune = on != 0;
    B->iter.reset();
}

}  // namespace bmg
