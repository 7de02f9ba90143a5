// Host-side solver context shared by the C ABI (bmg_capi.hip) and the
// batched solver (bmg_batch.hip): the level hierarchy with its device
// fields, the coarse factorisation, reduction buffers, captured graphs and
// the MG control flow that enqueues kernels on the context's stream.
#pragma once

#include "bmg.h"
#include "bmg_unet.h"
#include "bmg_internal.h"

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace bmg {

struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hipCheck(hipError_t e, const char* what, int line)
{
    if (e != hipSuccess)
        throw Error(BMG_ERR_HIP, std::string(what) + " (line " +
                                     std::to_string(line) + "): " +
                                     hipGetErrorString(e));
}
#define HIPCK(x) ::bmg::hipCheck((x), #x, __LINE__)

struct HostAxis
{
    int N = 0, R = 0;
    std::vector<double> v, cc, w, dl;
};

template <typename T>
T* devAlloc(size_t n)
{
    if (n == 0) return nullptr;
    void* p = nullptr;
    HIPCK(hipMalloc(&p, n*sizeof(T)));
    return static_cast<T*>(p);
}

template <typename T>
void devFree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T>
T* devUpload(const std::vector<T>& h)
{
    T* d = devAlloc<T>(h.size());
    HIPCK(hipMemcpy(d, h.data(), h.size()*sizeof(T), hipMemcpyHostToDevice));
    return d;
}

struct Level
{
    LevelDims L{};
    HostAxis ax[3];
    int opkind = OP_STENCIL;
    bool haveOp = false;
    // FP64 fields
    double* x[2] = {nullptr, nullptr};
    double* b[2] = {nullptr, nullptr};
    double* r[2] = {nullptr, nullptr};
    double* tmp[2] = {nullptr, nullptr};
    // FP32 fields (mixed precision)
    float* xf[2] = {nullptr, nullptr};
    float* bf[2] = {nullptr, nullptr};
    float* rf[2] = {nullptr, nullptr};
    float* tmpf[2] = {nullptr, nullptr};
    // operator
    double* a7[2] = {nullptr, nullptr};
    float* a7f[2] = {nullptr, nullptr};
    double* wf[2] = {nullptr, nullptr};
    _Float16* wff[2] = {nullptr, nullptr};   // FP16 faces, times wfs
    float wfs = 1.0f;
    double* dg[2] = {nullptr, nullptr};    // Euler time-term diagonal
    float* dgf[2] = {nullptr, nullptr};
    // metrics and transfer data
    double* dw[3] = {nullptr, nullptr, nullptr};
    double* ddl[3] = {nullptr, nullptr, nullptr};
    float* fw[3] = {nullptr, nullptr, nullptr};
    float* fdl[3] = {nullptr, nullptr, nullptr};
    double* ru[3] = {nullptr, nullptr, nullptr};
    double* pu[3] = {nullptr, nullptr, nullptr};
    float* ruf[3] = {nullptr, nullptr, nullptr};
    float* puf[3] = {nullptr, nullptr, nullptr};
    // snapshots of the slab planes 1 and N_d of every periodic axis d
    // (P_a1 x P_a2 values each), read by the slab smoother
    double* snap[3][2] = {{nullptr, nullptr}, {nullptr, nullptr},
                          {nullptr, nullptr}};
    float* snapf[3][2] = {{nullptr, nullptr}, {nullptr, nullptr},
                          {nullptr, nullptr}};

    size_t planeCells(int d) const
    {
        return size_t(d == 0 ? L.P[1] : L.P[0])*size_t(d == 2 ? L.P[1] : L.P[2]);
    }

    long cells() const { return long(L.N[0])*L.N[1]*L.N[2]; }

    void freeAll()
    {
        for (int c = 0; c < 2; c++)
        {
            if (rf[c] == bf[c]) rf[c] = nullptr;   // level-0 alias
            devFree(x[c]); devFree(b[c]); devFree(r[c]); devFree(tmp[c]);
            devFree(xf[c]); devFree(bf[c]); devFree(rf[c]); devFree(tmpf[c]);
            devFree(a7[c]); devFree(a7f[c]); devFree(wf[c]); devFree(wff[c]);
            devFree(dg[c]); devFree(dgf[c]);
        }
        for (int d = 0; d < 3; d++)
        {
            devFree(dw[d]); devFree(ddl[d]); devFree(fw[d]); devFree(fdl[d]);
            devFree(ru[d]); devFree(pu[d]); devFree(ruf[d]); devFree(puf[d]);
            for (int k = 0; k < 2; k++) { devFree(snap[d][k]); devFree(snapf[d][k]); }
        }
    }
};

template <typename T>
inline Pair<T> pr(T* const (&a)[2])
{
    return Pair<T>{{a[0], a[1]}};
}

template <typename T>
inline Pair<const T> cpr(T* const (&a)[2])
{
    return Pair<const T>{{a[0], a[1]}};
}


}  // namespace bmg

// The context is the opaque bmg_ctx of the C ABI, so it lives in the global
// namespace; its members use the solver's types.
using bmg::devAlloc;
using bmg::devFree;
using bmg::Level;
using bmg::LevelDims;
using bmg::OpView;
using bmg::Pair;
using bmg::Region;
using bmg::TransferView;

struct bmg_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    int bc[6];
    // level-0 boundary data: uniform values, or per-cell device arrays
    // (bcdev, owned) set by bmg_set_boundary_values
    bmg::BcData bcval{};
    double* bcdev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // level-0 ghost coefficients of a stored stencil saved before its
    // elimination (one per boundary cell of each non-periodic face)
    double* gsave[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // the next solve folds the level-0 boundary terms into its b (set with
    // the operator, and by bmg_renew_source_terms)
    bool foldPending = true;
    bmg_opts opts;
    std::vector<Level> lv;
    bool symmetric = true;
    bool eliminated = false;
    bool prepared = false;       // coarse factorisation valid
    int preparedCoarse = -1;
    // coarse LU
    double* dLU = nullptr;
    // inverse of the coarse matrix (64 x 64 column-major, zero-padded) for
    // the FP32 cycles, when the coarsest level has at most 64 cells
    double* dCoarseInv = nullptr;
    int* dPiv = nullptr;
    int nc = 0;
    // reductions
    double* partials = nullptr;
    size_t npartials = 0;
    double* dScalar = nullptr;
    double* hScalar = nullptr;   // pinned, mapped into the device
    double* hScalarDev = nullptr;   // device address of hScalar
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool mixedReady = false;
    // flexible CG state (level-0 FP64 fields) and the captured preconditioner
    double* kx[2] = {nullptr, nullptr};
    double* kb[2] = {nullptr, nullptr};
    double* kr[2] = {nullptr, nullptr};
    double* kp[2] = {nullptr, nullptr};
    double* kq[2] = {nullptr, nullptr};
    hipGraph_t pgraph = nullptr;
    hipGraphExec_t pgraphExec = nullptr;
    int pgraphMixed = -1;
    long pgraphGen = -1;
    // one flexible-CG iteration (operator, updates, preconditioner, next
    // direction) captured as a HIP graph
    hipGraph_t fgraph = nullptr;
    hipGraphExec_t fgraphExec = nullptr;
    int fgraphMode = -1;
    long fgraphGen = -1;
    // learned preconditioner: network (not owned), dense FP32 buffers of the
    // level-0 interior (input, coefficient channel, output) and the host
    // copy of the coefficient channel of the current operator
    bmg_unet* unet = nullptr;
    float* uIn = nullptr;
    float* uCoef = nullptr;
    float* uOut = nullptr;
    std::vector<float> coefChannel;
    double coefScale = 1.0;   // 10^mean(log10 lambda) of the cells
    bool coefUploaded = false;
    // flexible-CG deflation (opts.deflation): face-connected sets of
    // high-coefficient level-0 cells found when the operator is assembled
    // (host labels per interior cell, -1: none), and their device data,
    // built for operator generation deflOpGen
    std::vector<int> deflLabel;
    int deflM = 0;
    long opGen = 0, deflOpGen = -1;
    unsigned* dDeflCell = nullptr;   // component cells, by component
    int* dDeflCellComp = nullptr;
    long deflNCells = 0;
    unsigned* dDeflEntCell = nullptr;   // entries (A z_k)_i, by component
    double* dDeflEntVal = nullptr;
    // both lists cut into chunks that do not straddle components: chunk
    // ranges, and the chunks of component k at [off[k], off[k+1])
    int* dDeflCellBeg = nullptr;
    int* dDeflCellEnd = nullptr;
    int* dDeflCellChunkOff = nullptr;
    int deflCellChunks = 0;
    int* dDeflEntBeg = nullptr;
    int* dDeflEntEnd = nullptr;
    int* dDeflEntChunkOff = nullptr;
    int deflEntChunks = 0;
    double* dDeflPart = nullptr;     // chunk partial sums
    double* dDeflEinv = nullptr;     // (Z^T A Z)^-1, m x m
    // one MG iteration captured as a HIP graph and replayed
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    int graphMixed = -1;
    long gen = 0;          // bumped whenever device buffers are (re)allocated
    long graphGen = -1;
    bool useGraphs = true;
    // fused two-phase smoothing (launch_rbgs2) on levels without periodic
    // faces: 0 never (default: one launch per phase is faster on MI355X),
    // 1 on levels whose launch has at least 256 workgroups, 2 on every level
    // whose rows fit the LDS ring
    int fuseSmoothing = 0;
    // cycles on periodic levels: 1 (default) red-black phases over the
    // whole interior that keep the periodic ghosts current (wrapAxes) in the
    // FP32 cycles, 2 in the FP64 ones too, 0 nowhere (the reference's
    // interior phases, ghost fill and boundary slabs)
    int wrapSmoothing = 1;

    int nLevels() const { return int(lv.size()); }
    int coarseLevel() const { return nLevels() - opts.coarse_level - 1; }
    bool periodic(int f) const { return bc[f] == BMG_PERIODIC; }
    bool singular() const
    {
        if (!symmetric) return false;
        for (int f = 0; f < 6; f++)
            if (bc[f] == BMG_DIRICHLET) return false;
        return true;
    }
    // level-0 boundary terms to fold into b (eliminateGhosts, l == 0)
    bool inhomogeneous() const
    {
        for (int f = 0; f < 6; f++)
            if (!periodic(f) && (bcval.p[f] || bcval.v[f] != 0.0)) return true;
        return false;
    }
    bool anyPeriodic() const
    {
        for (int f = 0; f < 6; f++)
            if (periodic(f)) return true;
        return false;
    }

    // Periodic axes whose face ghosts the red-black phases of level l keep
    // current (Region::wrap): FP32 cycles (wrapSmoothing 1), or FP64 ones as
    // well (2; by default the FP64 path follows the reference's
    // interior-then-boundary-slab order bit for bit), and only when every
    // periodic axis has an even cell count, so that a ghost has the colour
    // of the cell it mirrors, of at least 8 (thinner periodic axes converge
    // better in the slab order, which sweeps the boundary planes one after
    // the other: measured 30 -> 59 cycles with two cells).  0: the
    // interior/slab smoother.
    template <typename T>
    int wrapAxes(int l) const
    {
        if (opts.smoother != BMG_RBGS || wrapSmoothing == 0 ||
            (!std::is_same<T, float>::value && wrapSmoothing < 2))
            return 0;
        int m = 0;
        for (int d = 0; d < 3; d++)
        {
            if (!periodic(2*d)) continue;
            if (lv[l].L.N[d] % 2 != 0 || lv[l].L.N[d] < 8) return 0;
            m |= 1 << d;
        }
        return m;
    }

    ~bmg_ctx()
    {
        for (auto& l : lv) l.freeAll();
        for (int f = 0; f < 6; f++) { devFree(bcdev[f]); devFree(gsave[f]); }
        devFree(dLU);
        devFree(dPiv);
        devFree(dCoarseInv);
        devFree(partials);
        devFree(dScalar);
        if (hScalar) (void)hipHostFree(hScalar);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        if (pgraphExec) (void)hipGraphExecDestroy(pgraphExec);
        if (pgraph) (void)hipGraphDestroy(pgraph);
        if (fgraphExec) (void)hipGraphExecDestroy(fgraphExec);
        if (fgraph) (void)hipGraphDestroy(fgraph);
        devFree(uIn);
        devFree(uCoef);
        devFree(uOut);
        freeDeflation();
        for (int c = 0; c < 2; c++)
        {
            devFree(kx[c]); devFree(kb[c]); devFree(kr[c]);
            devFree(kp[c]); devFree(kq[c]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ---------------------------------------------------------------- views
    template <typename T>
    OpView<T> op(int l) const;

    template <typename T>
    TransferView<T> transfer(int lf) const;

    template <typename T>
    Pair<T> X(int l);
    template <typename T>
    Pair<T> B(int l);
    template <typename T>
    Pair<T> Rr(int l);
    template <typename T>
    Pair<T> Tmp(int l);
    template <typename T>
    T* const (*Snap(int l))[2];

    // -------------------------------------------------------------- helpers
    double reduceFinal(int nblocks)
    {
        bmg::launch_reduce_final(partials, nblocks, dScalar, stream,
                                 hScalarDev);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(stream));
        return *hScalar;
    }

    Region interior(int l, int padding) const
    {
        Region rg;
        for (int d = 0; d < 3; d++)
        {
            const int padLo = periodic(2*d), padHi = periodic(2*d + 1);
            rg.S[d] = 1 + padding*padLo;
            rg.E[d] = lv[l].L.N[d] + 1 - padding*padHi;
        }
        return rg;
    }

    Region slab(int l, int f) const
    {
        Region rg;
        for (int d = 0; d < 3; d++) { rg.S[d] = 1; rg.E[d] = lv[l].L.N[d] + 1; }
        const int fd = f/2;
        if (f % 2 == 0) rg.E[fd] = rg.S[fd] + 1;
        else rg.S[fd] = rg.E[fd] - 1;
        return rg;
    }

    template <typename T>
    void smooth(int l, int sweeps, bool firstDone = false,
                bool firstZero = false);
    template <typename T>
    void runPhases(int l, const Region& rg, const std::vector<int>& ph,
                   bool firstZero);
    template <typename T>
    bool firstPhaseRewrites(int l) const;
    template <typename T>
    bool fusedLevel(int l) const;
    template <typename T>
    void smoothFused(int l, int sweeps);
    // buffers the smoothers use, allocated before any graph capture
    void ensureSmootherBuffers(bool mixed);
    // deflation of flexible CG
    void deflationLabels(const double* lx, const double* ly,
                         const double* lz);
    void ensureDeflation();
    bool deflating() const;
    void enqueueDeflate();
    void deflateStart();
    void freeDeflation()
    {
        devFree(dDeflCell); devFree(dDeflCellComp);
        devFree(dDeflEntCell); devFree(dDeflEntVal);
        devFree(dDeflCellBeg); devFree(dDeflCellEnd); devFree(dDeflCellChunkOff);
        devFree(dDeflEntBeg); devFree(dDeflEntEnd); devFree(dDeflEntChunkOff);
        devFree(dDeflPart); devFree(dDeflEinv);
        deflOpGen = -1;
    }
    template <typename T>
    void residual(int l, bool corrBcs, bool norm, int* nblocks);
    template <typename T>
    void restrictTo(int l);
    template <typename T>
    void prolongFrom(int l, int skipColour = -1);
    template <typename T>
    int postPhaseColour(int l) const;
    template <typename T>
    void coarseSolve();
    // zeroX: x of level l starts at 0 (every level below the finest; the
    // finest when a caller solves for a correction)
    template <typename T>
    void cycle(std::vector<int>& visits, int l, bool zeroX = false);

    void enqueueIteration(bool mixed);
    double iterate(bool mixed);
    void enqueuePrecondition(bool mixed);
    // FCG preconditioner modes: 0 FP64 cycle, 1 mixed cycle, 2 learned
    void enqueueFCGIteration(int mode);
    void fcgIteration(int mode);
    void enqueueLearned(const double* ssDev, double ssHost, Pair<double> z);
    void prepareLearned();
    void precondition(bool mixed);
    double normFactorOf(Pair<const double> x, Pair<const double> r,
                        Pair<const double> b);
    void runFCG(bool constMatrix, bmg_stats* st);
    void prepare(bool constMatrix);
    void prepareCoarse();
    void ensureMixed();
    void runSolve(bool constMatrix, bmg_stats* st);
    // runSolve in phases, each enqueued without synchronisation (and so
    // recordable for batched solves)
    bool prepareNeedsHost(bool constMatrix) const;
    // b[0] -= the level-0 boundary terms of eliminateGhosts, once per
    // assembled (or renewed) system
    void enqueueBoundaryTerms();
    // the FP32 correction fields zeroed, ghosts included (solve starts)
    void clearStaleCorrections();
    void enqueueSolveStart();
    void enqueueNormTerms(double y);
    void enqueueSolveEnd();
    // runFCG in phases (no learned preconditioner), likewise recordable
    void prepareFCG();
    void enqueueFCGStart();
    void enqueueFCGNormTerms(double y);
    void enqueueFCGRestart(bool mixed);
    void stageRestart(bool mixed);
    void finishRestart(bool mixed);
    void enqueueFCGConfirm();
    void enqueueFCGEnd();
};

template <> OpView<double> bmg_ctx::op<double>(int l) const;
template <> OpView<float> bmg_ctx::op<float>(int l) const;
template <> TransferView<double> bmg_ctx::transfer<double>(int lf) const;
template <> TransferView<float> bmg_ctx::transfer<float>(int lf) const;
template <> Pair<double> bmg_ctx::X<double>(int l);
template <> Pair<double> bmg_ctx::B<double>(int l);
template <> Pair<double> bmg_ctx::Rr<double>(int l);
template <> Pair<double> bmg_ctx::Tmp<double>(int l);
template <> Pair<float> bmg_ctx::X<float>(int l);
template <> Pair<float> bmg_ctx::B<float>(int l);
template <> Pair<float> bmg_ctx::Rr<float>(int l);
template <> Pair<float> bmg_ctx::Tmp<float>(int l);

struct bmg_batch;

namespace bmg {

// batched solves (bmg_batch.hip); throw Error on failure
void batch_create(bmg_batch** out, bmg_ctx* const* ctxs, int n);
void batch_solve(bmg_batch* b, double* const* db, double* const* dx,
                 bool constMatrix, bmg_stats* st);
void batch_info(const bmg_batch* b, int* launches, long* workgroups);
void batch_set_autotune(bmg_batch* b, int on);
void batch_tuning(const bmg_batch* b, int* launches, int* reshaped);
void batch_set_tile_shape(bmg_batch* b, int ws);
void batch_destroy(bmg_batch* b);

}  // namespace bmg

struct bmg_fft;

namespace bmg {

// FFT Poisson solver (bmg_fft.hip); throw Error on failure
void fft_create(bmg_fft** out, const bmg_grid* grid);
void fft_destroy(bmg_fft* f);
void fft_solve(bmg_fft* f, const double* b, double* x,
               const double* xold = nullptr, double dT = 0.0);
void fft_solve_device(bmg_fft* f, const double* db, double* dx,
                      const double* dxold = nullptr, double dT = 0.0);

}  // namespace bmg
