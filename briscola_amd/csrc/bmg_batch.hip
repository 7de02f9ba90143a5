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
// replayed; systems that have converged are dropped from the shares the
// kernels read from device memory, so the graph never changes during a
// solve.  Every merged launch is a persistent grid walking its shares; its
// size is chosen by timing the candidates once per plan.
#include "bmg_context.h"

#include <algorithm>
#include <array>
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
    std::vector<unsigned char> args0;  // the same as recorded
    int (*retile)(unsigned char*, int, unsigned*) = nullptr;
    int ws = 0;                        // tile shape picked (0: recorded)
    std::vector<BatchItem> items;      // every system's share
    std::vector<std::array<unsigned, 3>> grid0;   // recorded grids
    std::vector<int> sys;              // the system of each share
    unsigned total = 0;                // virtual workgroups, all systems
    size_t argsOff = 0, itemsOff = 0;  // offsets in the plan's regions
    unsigned grid = 0;                 // workgroups launched
};

}  // namespace

// A merged launch sequence with its device copy of argument structs and
// system shares.  The device buffer holds the argument structs, then the
// shares region: every launch's shares followed by one BatchState per
// launch.  Dropping systems rewrites the shares region only.
struct BatchPlan
{
    std::vector<Merged> launches;
    unsigned char* dev = nullptr;
    unsigned char* staging = nullptr;   // pinned copy of the shares region
    size_t sharesOff = 0, sharesBytes = 0, statesOff = 0;
    unsigned defGrid = 0;   // persistent grid before tuning
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (dev) (void)hipFree(dev);
        if (staging) (void)hipHostFree(staging);
    }

    // Merge of per-system sequences; a system's index is the index of its
    // sequence.  When every sequence's tags are non-decreasing, the
    // sequences are cut into segments of equal tag and the segments of each
    // tag are merged with each other, tags in increasing order; otherwise
    // the whole sequences are merged at once.  Within a merge, repeatedly
    // the kernel at the head of the most sequences (the longest remainder
    // breaking ties) is taken from all of them.  Launch grids start
    // persistent at `grid` workgroups.
    void build(std::vector<std::vector<LaunchRecord>>& seqs, unsigned grid)
    {
        const size_t n = seqs.size();
        bool tagged = true;
        std::vector<long> tags;
        for (const auto& q : seqs)
            for (size_t i = 0; i < q.size(); i++)
            {
                if (i > 0 && q[i].tag < q[i - 1].tag) tagged = false;
                tags.push_back(q[i].tag);
            }
        std::sort(tags.begin(), tags.end());
        tags.erase(std::unique(tags.begin(), tags.end()), tags.end());
        if (!tagged) tags.assign(1, 0);
        std::vector<size_t> head(n, 0), end(n, 0);
        for (long tag : tags)
        {
            for (size_t s = 0; s < n; s++)
            {
                end[s] = head[s];
                while (end[s] < seqs[s].size() &&
                       (!tagged || seqs[s][end[s]].tag == tag))
                    end[s]++;
            }
            mergeSegment(seqs, head, end);
        }
        layout(n, grid);
    }

    void mergeSegment(std::vector<std::vector<LaunchRecord>>& seqs,
                      std::vector<size_t>& head,
                      const std::vector<size_t>& end)
    {
        const size_t n = seqs.size();
        for (;;)
        {
            // kernels at the heads: number of sequences, longest remainder
            std::map<std::pair<const void*, size_t>,
                     std::pair<size_t, size_t>> heads;
            for (size_t s = 0; s < n; s++)
            {
                if (head[s] >= end[s]) continue;
                const LaunchRecord& r = seqs[s][head[s]];
                auto& h = heads[{r.kind, r.shm}];
                h.first++;
                h.second = std::max(h.second, end[s] - head[s]);
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
                if (head[s] >= end[s]) continue;
                const LaunchRecord& r = seqs[s][head[s]];
                if (r.kind != m.kind || r.shm != m.shm) continue;
                head[s]++;
                if (m.items.empty())
                {
                    m.many = r.many;
                    m.retile = r.retile;
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
                it.arg = int(m.items.size());
                m.items.push_back(it);
                m.grid0.push_back({r.grid[0], r.grid[1], r.grid[2]});
                m.sys.push_back(int(s));
                m.args.insert(m.args.end(), r.args.begin(), r.args.end());
                // each share starts at a multiple of 8 (its XCD order)
                m.total += unsigned(alignUp(wg, 8));
            }
            if (!m.items.empty())
            {
                m.args0 = m.args;
                launches.push_back(std::move(m));
            }
        }
    }

    // Device copies of the argument structs and shares (all systems on).
    void layout(size_t n, unsigned grid)
    {
        defGrid = grid;
        size_t bytes = 0;
        for (auto& m : launches)
        {
            bytes = alignUp(bytes, 256);
            m.argsOff = bytes;
            bytes += m.args.size();
            m.grid = std::min(grid, m.total);
        }
        sharesOff = alignUp(bytes, 256);
        size_t sh = 0;
        for (auto& m : launches)
        {
            sh = alignUp(sh, 16);
            m.itemsOff = sh;
            sh += m.items.size()*sizeof(BatchItem);
        }
        statesOff = alignUp(sh, 16);
        sharesBytes = statesOff + launches.size()*sizeof(BatchState);
        HIPCK(hipHostMalloc(&staging, std::max<size_t>(sharesBytes, 1)));
        HIPCK(hipMalloc(&dev, sharesOff + std::max<size_t>(sharesBytes, 1)));
        uploadAll(n);
    }

    // Argument structs and the shares of every system to the device.
    void uploadAll(size_t n)
    {
        std::vector<unsigned char> host(sharesOff, 0);
        for (auto& m : launches)
            std::memcpy(host.data() + m.argsOff, m.args.data(), m.args.size());
        HIPCK(hipMemcpy(dev, host.data(), host.size(), hipMemcpyHostToDevice));
        std::vector<char> all(n, 1);
        writeShares(all);
        HIPCK(hipMemcpy(dev + sharesOff, staging, sharesBytes,
                        hipMemcpyHostToDevice));
    }

    // Tile shape ws (3..6: rows of 2^ws cells per wavefront row; 0: the
    // recorded shapes) for every share of launch m whose kernel takes it;
    // the others keep their recorded tiles.  Host copies only.
    void applyShape(Merged& m, int ws)
    {
        m.args = m.args0;
        m.total = 0;
        for (size_t j = 0; j < m.items.size(); j++)
        {
            BatchItem& it = m.items[j];
            unsigned char* a = m.args.data() + size_t(it.arg)*m.argSize;
            unsigned g[3];
            if (ws != 0 && m.retile && m.retile(a, ws, g))
            {
                it.gx = g[0];
                it.gy = g[1];
                it.n = g[0]*g[1]*g[2];
            }
            else
            {
                it.gx = m.grid0[j][0];
                it.gy = m.grid0[j][1];
                it.n = m.grid0[j][0]*m.grid0[j][1]*m.grid0[j][2];
            }
            it.start = m.total;
            m.total += unsigned(alignUp(it.n, 8));
        }
        m.ws = ws;
    }

    // Shares of the systems that are on, renumbered from 0, into staging.
    void writeShares(const std::vector<char>& on)
    {
        BatchState* st = reinterpret_cast<BatchState*>(staging + statesOff);
        for (size_t i = 0; i < launches.size(); i++)
        {
            const Merged& m = launches[i];
            BatchItem* out = reinterpret_cast<BatchItem*>(staging + m.itemsOff);
            int k = 0;
            unsigned total = 0;
            for (size_t j = 0; j < m.items.size(); j++)
            {
                if (!on[m.sys[j]]) continue;
                BatchItem it = m.items[j];
                it.start = total;
                total += unsigned(alignUp(it.n, 8));
                out[k++] = it;
            }
            st[i].nitems = k;
            st[i].total = total;
        }
    }

    // Drops the systems that are off from every launch; the copy runs on
    // stream s before whatever is enqueued next.  The caller synchronises
    // s before the next call (the staging buffer is reused).
    void setActive(const std::vector<char>& on, hipStream_t s)
    {
        writeShares(on);
        HIPCK(hipMemcpyAsync(dev + sharesOff, staging, sharesBytes,
                             hipMemcpyHostToDevice, s));
    }

    void launch(size_t i, hipStream_t s) const
    {
        const Merged& m = launches[i];
        BatchLaunch b;
        b.args = dev + m.argsOff;
        b.items = reinterpret_cast<const BatchItem*>(dev + sharesOff +
                                                     m.itemsOff);
        b.state = reinterpret_cast<const BatchState*>(dev + sharesOff +
                                                      statesOff) + i;
        // a multiple of 8 workgroups (one group per XCD)
        b.grid = std::max(8u, m.grid/8*8);
        b.shm = m.shm;
        m.many(b, s);
    }

    void enqueue(hipStream_t s) const
    {
        for (size_t i = 0; i < launches.size(); i++) launch(i, s);
        HIPCK(hipGetLastError());
    }

    // Runs the sequence once, synchronously.
    void run(hipStream_t s) const
    {
        enqueue(s);
        HIPCK(hipStreamSynchronize(s));
    }

    // Runs the sequence once with an event after every launch; ms[i] is
    // launch i's time.
    void timeLaunches(hipStream_t s, std::vector<hipEvent_t>& ev,
                      std::vector<float>& ms) const
    {
        const size_t n = launches.size();
        HIPCK(hipEventRecord(ev[0], s));
        for (size_t i = 0; i < n; i++)
        {
            launch(i, s);
            HIPCK(hipEventRecord(ev[i + 1], s));
        }
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(s));
        ms.assign(n, 0.0f);
        for (size_t i = 0; i < n; i++)
            HIPCK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
    }

    // Picks, for every launch, first the tile shape of its shares (rows of
    // 64, 32, 16 or 8 cells per wavefront row, or the recorded shapes; for
    // the kernels that take several: red-black phases, residuals without a
    // norm, restriction, prolongation), then its persistent grid, 4, 8, 16
    // or 32 workgroups per CU (never more than its workgroups), each by the
    // fastest of timed runs of the whole sequence (the second of two).
    // Every launch that takes several tile shapes to shape ws (3..6).
    void fixShape(int ws, size_t nsys)
    {
        for (auto& m : launches)
            if (m.retile)
            {
                applyShape(m, ws);
                m.grid = std::min(defGrid, m.total);
            }
        uploadAll(nsys);
    }

    void autotune(hipStream_t s, int cus, size_t nsys, int fixedShape = 0)
    {
        const size_t n = launches.size();
        std::vector<hipEvent_t> ev(n + 1);
        for (auto& e : ev) HIPCK(hipEventCreate(&e));
        std::vector<float> ms;
        try
        {
            bool shapes = false;
            for (const auto& m : launches) shapes = shapes || m.retile;
            if (fixedShape != 0)
                fixShape(fixedShape, nsys);
            else if (shapes)
            {
                const int cand[5] = {0, 6, 5, 4, 3};
                std::vector<float> best(n, 1e30f);
                std::vector<int> pick(n, 0);
                for (int rep = 0; rep < 2; rep++)
                    for (int c : cand)
                    {
                        for (auto& m : launches)
                            if (m.retile)
                            {
                                applyShape(m, c);
                                m.grid = std::min(defGrid, m.total);
                            }
                        uploadAll(nsys);
                        timeLaunches(s, ev, ms);
                        for (size_t i = 0; i < n; i++)
                            if (rep > 0 && launches[i].retile && ms[i] < best[i])
                            {
                                best[i] = ms[i];
                                pick[i] = c;
                            }
                    }
                for (size_t i = 0; i < n; i++)
                    if (launches[i].retile)
                    {
                        applyShape(launches[i], pick[i]);
                        launches[i].grid = std::min(defGrid, launches[i].total);
                    }
                uploadAll(nsys);
            }
            const unsigned cand[4] = {unsigned(4*cus), unsigned(8*cus),
                                      unsigned(16*cus), unsigned(32*cus)};
            std::vector<float> best(n, 1e30f);
            std::vector<unsigned> pick(n, 0);
            for (int rep = 0; rep < 2; rep++)
                for (unsigned c : cand)
                {
                    for (auto& m : launches) m.grid = std::min(c, m.total);
                    timeLaunches(s, ev, ms);
                    for (size_t i = 0; i < n; i++)
                        if (rep > 0 && ms[i] < best[i])
                        {
                            best[i] = ms[i];
                            pick[i] = launches[i].grid;
                        }
                }
            for (size_t i = 0; i < n; i++) launches[i].grid = pick[i];
        }
        catch (...)
        {
            for (auto& e : ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
    }

    // Launches whose tile shape the tuner changed from the recorded one.
    int reshaped() const
    {
        int k = 0;
        for (const auto& m : launches) k += m.ws != 0;
        return k;
    }

    void capture(hipStream_t s)
    {
        HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try
        {
            enqueue(s);
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
    std::unique_ptr<BatchPlan> iter;
    std::vector<long> iterGen;   // context generations the plan was built on
    std::vector<int> iterMixed;
    // flexible CG: the iteration (captured), the restart p = M r and the
    // confirmation on the true residual
    std::unique_ptr<BatchPlan> fcgIter, fcgRestart, fcgConfirm;
    std::vector<long> fcgGen;
    std::vector<int> fcgMixed;
    bool lastFcg = false;   // the last solve ran flexible CG
    bool tune = true;
    int shape = 0;   // tile shape of every launch that takes one (0: tuned)
    int cus = 256;

    ~bmg_batch()
    {
        iter.reset();
        fcgIter.reset();
        fcgRestart.reset();
        fcgConfirm.reset();
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
    p->build(seqs, unsigned(16*B->cus));
    return p;
}

bool mixedOf(const bmg_ctx* c) { return c->opts.precision == BMG_MIXED; }

// Flexible CG of every context as bmg_ctx::runFCG runs it (no learned
// preconditioner): every phase between two host decisions is recorded and
// run as merged launches over the systems it concerns -- the restart
// p = M r for the systems that need one, one captured iteration for the
// systems still iterating, the confirmation on the true residual for the
// systems whose updated residual met the tolerance.
void batchFCG(bmg_batch* B, double* const* db, double* const* dx,
              bool constMatrix, const std::vector<char>& hostPrepared,
              std::vector<int>& iter, std::vector<double>& init,
              std::vector<double>& cur)
{
    const size_t n = B->ctx.size();
    bool rebuild = !B->fcgIter;
    for (size_t s = 0; s < n && !rebuild; s++)
        rebuild = B->fcgGen[s] != B->ctx[s]->gen ||
                  B->fcgMixed[s] != int(mixedOf(B->ctx[s]));
    if (rebuild)
    {
        B->fcgIter = recordPlan(B, [](bmg_ctx* c, size_t) {
            c->enqueueFCGIteration(mixedOf(c) ? 1 : 0);
        });
        B->fcgRestart = recordPlan(B, [](bmg_ctx* c, size_t) {
            c->enqueueFCGRestart(mixedOf(c));
        });
        B->fcgConfirm = recordPlan(B, [](bmg_ctx* c, size_t) {
            c->enqueueFCGConfirm();
        });
        B->fcgGen.resize(n);
        B->fcgMixed.resize(n);
        for (size_t s = 0; s < n; s++)
        {
            B->fcgGen[s] = B->ctx[s]->gen;
            B->fcgMixed[s] = int(mixedOf(B->ctx[s]));
        }
        // timing runs before the solve's fields are initialised below
        if (B->tune)
            B->fcgIter->autotune(B->stream, B->cus, n, B->shape);
        else if (B->shape != 0)
            B->fcgIter->fixShape(B->shape, n);
        B->fcgIter->capture(B->stream);
    }
    // copies of x and b, initial residual, sum of x
    recordPlan(B, [&](bmg_ctx* c, size_t s) {
        Level& q = c->lv[0];
        launch_ref_to_split<double>(q.L, db[s], pr(q.b), c->stream);
        launch_ref_to_split<double>(q.L, dx[s], pr(q.x), c->stream);
        if (!hostPrepared[s]) c->prepare(constMatrix);
        c->enqueueFCGStart();
    })->run(B->stream);
    std::vector<double> nf(n), ss(n);
    for (size_t s = 0; s < n; s++) ss[s] = B->ctx[s]->hScalar[0];
    recordPlan(B, [&](bmg_ctx* c, size_t) {
        c->enqueueFCGNormTerms(c->hScalar[1]/double(c->lv[0].cells()));
    })->run(B->stream);
    auto check = [&](size_t s, double v) -> int
    {
        const bmg_opts& o = B->ctx[s]->opts;
        return (init[s] > 0.0 && v > o.tolerance &&
                v/(init[s] + 1e-20) > o.rel_tol) ? 0 : 1;
    };
    std::vector<int> conv(n, 0);
    std::vector<char> restart(n, 1), done(n, 0);
    for (size_t s = 0; s < n; s++)
    {
        bmg_ctx* c = B->ctx[s];
        const double cells = double(c->lv[0].cells());
        nf[s] = std::max(std::sqrt(c->hScalar[2]), cells*1e-12);
        init[s] = std::sqrt(ss[s])/nf[s];
        if (!std::isfinite(init[s]))
            throw Error(BMG_ERR_NUMERIC, "initial residual is not finite");
        cur[s] = init[s];
        conv[s] = check(s, cur[s]);
    }
    std::vector<char> go(n), rs(n), cf(n), lastGo;
    for (;;)
    {
        size_t ngo = 0, nrs = 0;
        for (size_t s = 0; s < n; s++)
        {
            const bmg_opts& o = B->ctx[s]->opts;
            go[s] = !done[s] && ((iter[s] < o.max_iter && conv[s] == 0) ||
                                 iter[s] < o.min_iter);
            rs[s] = go[s] && restart[s];
            ngo += go[s];
            nrs += rs[s];
        }
        if (ngo == 0) break;
        if (nrs > 0)
        {
            B->fcgRestart->setActive(rs, B->stream);
            B->fcgRestart->run(B->stream);
            for (size_t s = 0; s < n; s++)
                if (rs[s]) restart[s] = 0;
        }
        if (go != lastGo)
        {
            B->fcgIter->setActive(go, B->stream);
            lastGo = go;
        }
        HIPCK(hipGraphLaunch(B->fcgIter->exec, B->stream));
        HIPCK(hipStreamSynchronize(B->stream));
        size_t ncf = 0;
        for (size_t s = 0; s < n; s++)
        {
            cf[s] = 0;
            if (!go[s]) continue;
            bmg_ctx* c = B->ctx[s];
            const double pq = c->hScalar[0];
            if (!(pq != 0.0) || !std::isfinite(pq))
                throw Error(BMG_ERR_NUMERIC, "flexible CG breakdown (p.Ap = 0)");
            cur[s] = std::sqrt(c->hScalar[2])/nf[s];
            if (!std::isfinite(cur[s]))
                throw Error(BMG_ERR_NUMERIC, "residual norm is not finite");
            iter[s]++;
            conv[s] = check(s, cur[s]);
            if (conv[s] && iter[s] >= c->opts.min_iter)
            {
                cf[s] = 1;
                ncf++;
            }
        }
        if (ncf == 0) continue;
        // confirm on the true residual (with the boundary update)
        B->fcgConfirm->setActive(cf, B->stream);
        B->fcgConfirm->run(B->stream);
        for (size_t s = 0; s < n; s++)
        {
            if (!cf[s]) continue;
            cur[s] = std::sqrt(B->ctx[s]->hScalar[0])/nf[s];
            conv[s] = check(s, cur[s]);
            if (conv[s]) done[s] = 1;
            else restart[s] = 1;
        }
    }
    // x and b back, eliminated ghosts, results in the reference layout
    recordPlan(B, [&](bmg_ctx* c, size_t s) {
        Level& q = c->lv[0];
        c->enqueueFCGEnd();
        launch_split_to_ref<double>(q.L, cpr(q.x), dx[s], c->stream);
        launch_split_to_ref<double>(q.L, cpr(q.b), db[s], c->stream);
    })->run(B->stream);
}

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
        if (c->opts.krylov == BMG_KRYLOV_FCG_LEARNED)
            throw Error(BMG_ERR_ARG, "the learned preconditioner has no "
                                     "batched form");
        if (c->opts.deflation)
            throw Error(BMG_ERR_ARG, "deflated flexible CG has no batched "
                                     "form");
        if (c->opts.krylov != B->ctx[0]->opts.krylov)
            throw Error(BMG_ERR_ARG, "every system of a batch uses the same "
                                     "outer iteration (MG or flexible CG)");
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
            if (c->opts.krylov == BMG_KRYLOV_FCG)
            {
                c->prepareFCG();
                HIPCK(hipStreamSynchronize(c->stream));
            }
        }
        std::vector<double> init(n), cur(n);
        std::vector<int> iter(n, 0);
        B->lastFcg = B->ctx[0]->opts.krylov == BMG_KRYLOV_FCG;
        if (B->lastFcg)
            batchFCG(B, db, dx, constMatrix, hostPrepared, iter, init, cur);
        else
        {
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
                if (B->tune)
                    B->iter->autotune(B->stream, B->cus, n, B->shape);
                else if (B->shape != 0)
                    B->iter->fixShape(B->shape, n);
                B->iter->capture(B->stream);
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
            std::vector<double> nf(n);
            std::vector<int> conv(n, 0);
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
            // the captured iteration walks the shares of the systems still
            // iterating; they are rewritten whenever that set changes
            std::vector<char> active(n, 1), last;
            for (;;)
            {
                size_t nactive = 0;
                for (size_t s = 0; s < n; s++)
                {
                    const bmg_opts& o = B->ctx[s]->opts;
                    const bool go = (iter[s] < o.max_iter && conv[s] == 0) ||
                                    iter[s] < o.min_iter;
                    active[s] = go;
                    nactive += go;
                }
                if (nactive == 0) break;
                if (active != last)
                {
                    B->iter->setActive(active, B->stream);
                    last = active;
                }
                HIPCK(hipGraphLaunch(B->iter->exec, B->stream));
                HIPCK(hipStreamSynchronize(B->stream));
                for (size_t s = 0; s < n; s++)
                {
                    if (!active[s]) continue;
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
        }
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
    const BatchPlan* p = B->lastFcg ? B->fcgIter.get() : B->iter.get();
    if (launches) *launches = p ? int(p->launches.size()) : 0;
    if (workgroups) *workgroups = p ? long(p->workgroups()) : 0;
}

void batch_tuning(const bmg_batch* B, int* launches, int* reshaped)
{
    if (!B) throw Error(BMG_ERR_ARG, "null batch");
    const BatchPlan* p = B->lastFcg ? B->fcgIter.get() : B->iter.get();
    if (launches) *launches = p ? int(p->launches.size()) : 0;
    if (reshaped) *reshaped = p ? p->reshaped() : 0;
}

void batch_set_tile_shape(bmg_batch* B, int ws)
{
    if (!B) throw Error(BMG_ERR_ARG, "null batch");
    if (ws != 0 && (ws < 3 || ws > 6))
        throw Error(BMG_ERR_ARG, "tile shape must be 0 or 3..6");
    B->shape = ws;
    B->iter.reset();
    B->fcgIter.reset();
}

void batch_destroy(bmg_batch* B) { delete B; }

void batch_set_autotune(bmg_batch* B, int on)
{
    if (!B) throw Error(BMG_ERR_ARG, "null batch");
    B->tune = on != 0;
    B->iter.reset();
    B->fcgIter.reset();
}

}  // namespace bmg
