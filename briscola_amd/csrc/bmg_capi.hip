// Host side of the MI355X multigrid solver: level hierarchy, metrics,
// transfer weights, the MG::solve / Solve / cycle control flow and the C ABI
// declared in include/bmg.h.  All numerical work runs in the HIP kernels of
// bmg_kernels.hip on the context's stream.
#include "bmg.h"
#include "bmg_unet.h"
#include "unet_internal.h"
#include "bmg_internal.h"
#include "bmg_context.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

using namespace bmg;

namespace {

thread_local std::string g_error = "";


// Levels with more cells than this keep a matrix-free / face-coefficient
// operator; smaller ones (and the pinned coarsest level) store the stencil.
constexpr long MATERIALIZE_CELLS = 32768;

// utils.H:1010-1021
double trimPrecision(double s)
{
    if (s != 0.0)
    {
        const double factor =
            std::pow(10.0, 14 - 1 - static_cast<int>(std::log10(std::fabs(s))));
        s = std::round(s*factor)/factor;
    }
    return s;
}

// vertexVector.C:183-188
double round12(double u) { return std::round(u/1e-12)*1e-12; }

// hexa.H:120-128, one axis of a rectilinear cell
double hexaCentre(double lo, double hi, int axis)
{
    static const int pat[3][8] = {{0,1,0,1,0,1,0,1}, {0,0,1,1,0,0,1,1},
                                  {0,0,0,0,1,1,1,1}};
    double c = (pat[axis][0] ? hi : lo)/8.0;
    for (int i = 1; i < 8; i++) c += (pat[axis][i] ? hi : lo)/8.0;
    return c;
}

}  // namespace


template <>
OpView<double> bmg_ctx::op<double>(int l) const
{
    const Level& q = lv[l];
    OpView<double> o{};
    o.kind = q.opkind;
    for (int c = 0; c < 2; c++)
    {
        o.a7[c] = q.a7[c]; o.wf[c] = q.wf[c]; o.dg[c] = q.dg[c];
    }
    o.fs = 1.0;
    for (int d = 0; d < 3; d++) { o.w[d] = q.dw[d]; o.dl[d] = q.ddl[d]; }
    return o;
}

template <>
OpView<float> bmg_ctx::op<float>(int l) const
{
    const Level& q = lv[l];
    OpView<float> o{};
    o.kind = q.opkind;
    for (int c = 0; c < 2; c++)
    {
        o.a7[c] = q.a7f[c]; o.wf[c] = q.wff[c]; o.dg[c] = q.dgf[c];
    }
    o.fs = q.wfs;
    for (int d = 0; d < 3; d++) { o.w[d] = q.fw[d]; o.dl[d] = q.fdl[d]; }
    return o;
}

template <>
TransferView<double> bmg_ctx::transfer<double>(int lf) const
{
    TransferView<double> t{};
    for (int d = 0; d < 3; d++)
    {
        t.ru[d] = lv[lf + 1].ru[d];
        t.pu[d] = lv[lf].pu[d];
        t.fw[d] = lv[lf].dw[d];
        t.cw[d] = lv[lf + 1].dw[d];
    }
    return t;
}

template <>
TransferView<float> bmg_ctx::transfer<float>(int lf) const
{
    TransferView<float> t{};
    for (int d = 0; d < 3; d++)
    {
        t.ru[d] = lv[lf + 1].ruf[d];
        t.pu[d] = lv[lf].puf[d];
        t.fw[d] = lv[lf].fw[d];
        t.cw[d] = lv[lf + 1].fw[d];
    }
    return t;
}

template <> Pair<double> bmg_ctx::X<double>(int l) { return pr(lv[l].x); }
template <> Pair<double> bmg_ctx::B<double>(int l) { return pr(lv[l].b); }
template <> Pair<double> bmg_ctx::Rr<double>(int l) { return pr(lv[l].r); }
template <> Pair<double> bmg_ctx::Tmp<double>(int l)
{
    Level& q = lv[l];
    if (!q.tmp[0])
    {
        for (int c = 0; c < 2; c++) q.tmp[c] = devAlloc<double>(q.L.ncol);
        gen++;
    }
    return pr(q.tmp);
}
template <> double* const (*bmg_ctx::Snap<double>(int l))[2]
{
    return lv[l].snap;
}
template <> float* const (*bmg_ctx::Snap<float>(int l))[2]
{
    return lv[l].snapf;
}
template <> Pair<float> bmg_ctx::X<float>(int l) { return pr(lv[l].xf); }
template <> Pair<float> bmg_ctx::B<float>(int l) { return pr(lv[l].bf); }
template <> Pair<float> bmg_ctx::Rr<float>(int l) { return pr(lv[l].rf); }
template <> Pair<float> bmg_ctx::Tmp<float>(int l)
{
    Level& q = lv[l];
    if (!q.tmpf[0])
    {
        for (int c = 0; c < 2; c++) q.tmpf[c] = devAlloc<float>(q.L.ncol);
        gen++;
    }
    return pr(q.tmpf);
}

template <typename T>
static Pair<const T> cst(Pair<T> p)
{
    return Pair<const T>{{p.c[0], p.c[1]}};
}

// The colours of the four counter-relative phases of a region
// (rbgsSmoother.C:65-183): forward even = colour of S, forward odd = the
// other, backward even/odd anchored at E-1; a phase equal to the one before
// is dropped.  Returns the number of phases.
static int phase_colours(const Region& rg, int out[4])
{
    const int Ss = rg.S[0] + rg.S[1] + rg.S[2];
    const int Es = rg.E[0] + rg.E[1] + rg.E[2];
    const int c4[4] = {Ss & 1, (Ss + 1) & 1, (Es + 1) & 1, Es & 1};
    int n = 0;
    for (int i = 0; i < 4; i++)
        if (n == 0 || out[n - 1] != c4[i]) out[n++] = c4[i];
    return n;
}

template <typename T>
bool bmg_ctx::fusedLevel(int l) const
{
    if (fuseSmoothing == 0 || opts.smoother != BMG_RBGS || anyPeriodic() ||
        launch_recorder())
        return false;
    const int tiles = fused_rbgs_tiles<T>(lv[l].L);
    return tiles > 0 && (fuseSmoothing == 2 || tiles >= 256);
}

// Smoothing of a level without periodic faces, whose sweeps are interior
// phases only: the phase sequence of all sweeps (a phase equal to the one
// before skipped, as in smooth) is run as fused pairs (c1, c2) plus a last
// single c1 phase when the count is odd.  Colour c2 moves between x and the
// level's spare buffer with every pair and is copied back at the end when
// it finishes in the spare buffer.
template <typename T>
void bmg_ctx::smoothFused(int l, int sweeps)
{
    Level& q = lv[l];
    const OpView<T> o = op<T>(l);
    Pair<T> x = X<T>(l);
    Pair<T> bb = B<T>(l);
    const Region rg = interior(l, 1);
    std::vector<int> ph;
    for (int s = 0; s < sweeps; s++)
    {
        int cols[4];
        const int nc = phase_colours(rg, cols);
        for (int i = 0; i < nc; i++)
            if (ph.empty() || ph.back() != cols[i]) ph.push_back(cols[i]);
    }
    if (ph.empty()) return;
    T* spare = Tmp<T>(l).c[0];
    const int c1 = ph[0], c2 = c1 ^ 1;
    bool inSpare = false;
    size_t i = 0;
    while (i < ph.size())
    {
        T* cur2 = inSpare ? spare : x.c[c2];
        if (i + 1 < ph.size())
        {
            T* next2 = inSpare ? x.c[c2] : spare;
            launch_rbgs2(q.L, o, x.c[c1], cur2, next2, bb.c[c1], bb.c[c2], c1,
                         rg, stream);
            inSpare = !inSpare;
            i += 2;
        }
        else
        {
            Pair<T> xs;
            xs.c[c1] = x.c[c1];
            xs.c[c2] = cur2;
            launch_rbgs(q.L, o, xs, cst(bb), c1, rg, stream);
            i += 1;
        }
    }
    if (inSpare)
        HIPCK(hipMemcpyAsync(x.c[c2], spare, size_t(q.L.ncol)*sizeof(T),
                             hipMemcpyDeviceToDevice, stream));
}

void bmg_ctx::ensureSmootherBuffers(bool mixed)
{
    for (int l = 0; l < nLevels(); l++)
    {
        const bool need = opts.smoother == BMG_JAC ||
            (mixed ? fusedLevel<float>(l) : fusedLevel<double>(l));
        if (need) mixed ? (void)Tmp<float>(l) : (void)Tmp<double>(l);
    }
}

// rbgsSmoother::Smooth (rbgsSmoother.C:376-437) / jacSmoother::Smooth.  The
// four counter-relative phases of every region (rbgsSmoother.C:65-183) are
// turned into colours of the padded index: forward even = colour of S,
// forward odd = the other, backward even/odd anchored at E-1.  A phase that
// repeats the phase right before it on the same region is skipped: it reads
// only the other colour, which did not change, so it would rewrite the same
// values.  The boundary slabs of each periodic face pair run as one
// temporally blocked launch (k_rbgs_slabs).
// The red-black phases ph (colours in order) over region rg of level l, the
// first from x = 0 without reading its neighbours when firstZero: on levels
// of at most 16^3 cells one launch of one workgroup for all of them (the
// same bits), else one launch each.  Recorded (batched) sequences keep one
// launch each.
template <typename T>
void bmg_ctx::runPhases(int l, const Region& rg, const std::vector<int>& ph,
                        bool firstZero)
{
    Level& q = lv[l];
    const OpView<T> o = op<T>(l);
    Pair<T> x = X<T>(l);
    Pair<const T> b = cst(B<T>(l));
    if (ph.empty()) return;
    if (!launch_recorder() && q.cells() <= 4096 && ph.size() <= 8)
    {
        launch_rbgs_chain(q.L, o, x, b, rg, ph.data(), int(ph.size()),
                          firstZero ? 1 : 0, stream);
        return;
    }
    for (size_t i = 0; i < ph.size(); i++)
    {
        Region r = rg;
        r.zeroNb = i == 0 && firstZero;
        launch_rbgs(q.L, o, x, b, ph[i], r, stream);
    }
}

template <typename T>
void bmg_ctx::smooth(int l, int sweeps, bool firstDone, bool firstZero)
{
    Level& q = lv[l];
    const OpView<T> o = op<T>(l);
    Pair<T> x = X<T>(l);
    Pair<const T> b = cst(B<T>(l));
    if (opts.smoother == BMG_LEXGS)
    {
        // lexgsSmoother::Smooth (lexgsSmoother.C:256-317): interior region,
        // periodic ghosts, then the width-1 slab of every periodic face
        for (int s = 0; s < sweeps; s++)
        {
            launch_lexgs(q.L, o, x, b, interior(l, 1), stream);
            if (anyPeriodic())
            {
                launch_bc_fill(q.L, o, x, 0, l == 0, &bcval, stream);
                for (int f = 0; f < 6; f++)
                    if (periodic(f))
                        launch_lexgs(q.L, o, x, b, slab(l, f), stream);
            }
        }
        return;
    }
    if (opts.smoother == BMG_JAC)
    {
        Pair<T> tmp = Tmp<T>(l);
        for (int s = 0; s < sweeps; s++)
        {
            launch_jacobi(q.L, o, x, b, tmp, interior(l, 1), stream);
            if (anyPeriodic())
            {
                launch_bc_fill(q.L, o, x, 0, l == 0, &bcval, stream);
                for (int f = 0; f < 6; f++)
                    if (periodic(f))
                        launch_jacobi(q.L, o, x, b, tmp, slab(l, f), stream);
            }
        }
        return;
    }
    if (fusedLevel<T>(l))
    {
        smoothFused<T>(l, sweeps);
        return;
    }
    auto colours = phase_colours;
    if (const int wrap = wrapAxes<T>(l))
    {
        // even periodic axes: phases over the whole interior that write the
        // periodic ghosts with their cells; the ghosts are made current once
        // on entry unless launch_zero_phase left them so
        Region rg = interior(l, 0);
        rg.wrap = wrap;
        // with firstZero the first phase reads no neighbour and writes the
        // ghosts of its own colour, which is all the second one reads
        if (!firstDone && !firstZero)
            launch_bc_fill(q.L, o, x, 0, l == 0, &bcval, stream);
        std::vector<int> ph;
        int prev = -1;
        for (int s = 0; s < sweeps; s++)
        {
            int cols[4];
            const int nc = colours(rg, cols);
            for (int i = 0; i < nc; i++)
            {
                if (s == 0 && i == 0 && firstDone)
                {
                    prev = cols[0];
                    continue;
                }
                if (cols[i] != prev) ph.push_back(cols[i]);
                prev = cols[i];
            }
        }
        runPhases<T>(l, rg, ph, firstZero);
        return;
    }
    if (!anyPeriodic())
    {
        // no periodic face: the phases of every sweep over the interior
        const Region rg = interior(l, 1);
        std::vector<int> ph;
        int prev = -1;
        for (int s = 0; s < sweeps; s++)
        {
            int cols[4];
            const int nc = colours(rg, cols);
            for (int i = 0; i < nc; i++)
            {
                const bool done = s == 0 && i == 0 && firstDone;
                if (!done && cols[i] != prev) ph.push_back(cols[i]);
                prev = cols[i];
            }
        }
        runPhases<T>(l, rg, ph, firstZero);
        return;
    }
    // the previous phase launched on this level: a phase that repeats it on
    // the same region would rewrite the same values and is skipped
    int prevColour = -1;
    for (int s = 0; s < sweeps; s++)
    {
        // interior region, padded on the periodic faces
        const Region rg = interior(l, 1);
        int cols[4];
        const int nc = colours(rg, cols);
        // the first phase already run from x = 0 (launch_zero_phase)
        auto done = [&](int i) { return s == 0 && i == 0 && firstDone; };
        // periodic ghosts, plus snapshots of the slab planes: filled in the
        // launch of the sweep's last interior phase (the fill only reads
        // slab cells, which the interior phases do not write)
        T* const (*sn)[2] = anyPeriodic() ? Snap<T>(l) : nullptr;
        int last = -1;
        for (int i = 0, pc = prevColour; i < nc; i++)
        {
            if (!done(i) && cols[i] != pc) last = i;
            pc = cols[i];
        }
        for (int i = 0; i < nc; i++)
        {
            if (!done(i) && cols[i] != prevColour)
            {
                Region r = rg;
                r.zeroNb = s == 0 && i == 0 && firstZero;
                launch_rbgs_fill(q.L, o, x, b, cols[i], r, l == 0, &bcval,
                                 i == last ? sn : nullptr, stream);
            }
            prevColour = cols[i];
        }
        if (!anyPeriodic()) continue;
        if (last < 0)
            launch_bc_fill_snap(q.L, o, x, 0, l == 0, &bcval, sn, stream);
        // width-1 slabs of the periodic faces, face pairs in order: one
        // launch per pair when its two planes are independent, else one per
        // plane
        struct SlabLaunch { int d, slot, nslab, ncol; int cl[8]; };
        std::vector<SlabLaunch> seq;
        for (int d = 0; d < 3; d++)
        {
            if (!periodic(2*d)) continue;
            SlabLaunch s0{d, 0, 2, 0, {}};
            const int n0 = colours(slab(l, 2*d), s0.cl);
            const int n1 = colours(slab(l, 2*d + 1), s0.cl + 4);
            if (q.L.N[d] >= 3 && n0 == n1)
            {
                s0.ncol = n0;
                seq.push_back(s0);
            }
            else
            {
                SlabLaunch s1 = s0;
                s0.nslab = 1;
                s0.ncol = n0;
                s1.slot = 1;
                s1.nslab = 1;
                s1.ncol = n1;
                for (int i = 0; i < 4; i++) s1.cl[i] = s0.cl[4 + i];
                seq.push_back(s0);
                seq.push_back(s1);
            }
        }
        for (size_t i = 0; i < seq.size(); i++)
        {
            const SlabLaunch& sl = seq[i];
            const int d = sl.d;
            const int planes[2] = {1, q.L.N[d]};
            SlabIO<T> io{};
            io.src[0] = sn[d][sl.slot];
            io.src[1] = sn[d][1];
            // later launches whose planes meet this one
            for (size_t j = i + 1; j < seq.size(); j++)
                for (int k = 0; k < seq[j].nslab; k++)
                {
                    const int dj = seq[j].d, sj = seq[j].slot + k;
                    const int pj = sj ? q.L.N[dj] : 1;
                    if (dj == d && pj != planes[sl.slot]) continue;
                    io.tptr[io.ntarget] = sn[dj][sj];
                    io.tdim[io.ntarget] = dj;
                    io.tplane[io.ntarget] = pj;
                    io.ntarget++;
                }
            launch_rbgs_slabs(q.L, o, x, b, d, planes + sl.slot, sl.nslab,
                              sl.cl, sl.ncol, io, stream);
        }
        prevColour = -1;
    }
}

// linearSystem::residual (linearSystemResidual.C:280-332).  With BC
// correction the periodic ghosts are refreshed first: the interior part of
// the reference's split never reads them, so the values are the same.
template <typename T>
void bmg_ctx::residual(int l, bool corrBcs, bool norm, int* nblocks)
{
    Level& q = lv[l];
    const OpView<T> o = op<T>(l);
    if (corrBcs && anyPeriodic())
        launch_bc_fill(q.L, o, X<T>(l), 0, l == 0, &bcval, stream);
    launch_residual(q.L, o, cst(X<T>(l)), cst(B<T>(l)), Rr<T>(l),
                    norm ? partials : nullptr, nblocks, stream);
}

template <typename T>
void bmg_ctx::restrictTo(int l)
{
    launch_restrict(lv[l].L, lv[l + 1].L, transfer<T>(l), cst(Rr<T>(l)),
                    B<T>(l + 1), stream);
}

template <typename T>
void bmg_ctx::prolongFrom(int l, int skipColour)
{
    // the coarse ghosts are used, so correct them first
    // (linearProlongationScheme.C:128)
    launch_bc_fill(lv[l + 1].L, op<T>(l + 1), X<T>(l + 1), 1, (l + 1) == 0,
                   &bcval, stream);
    launch_prolong(lv[l].L, lv[l + 1].L, transfer<T>(l), cst(X<T>(l + 1)),
                   X<T>(l), stream, skipColour);
}

// The first sweep of red-black smoothing on level l covers every interior
// cell with at least two phases -- no periodic faces, or the wrapped phases --
// so the second phase rewrites every cell of the colour the first one reads.
template <typename T>
bool bmg_ctx::firstPhaseRewrites(int l) const
{
    if (opts.smoother != BMG_RBGS || fusedLevel<T>(l)) return false;
    const int wrap = wrapAxes<T>(l);
    if (anyPeriodic() && !wrap) return false;
    int cols[4];
    return phase_colours(interior(l, wrap ? 0 : 1), cols) >= 2;
}

// Colour of the first red-black phase of level l's post-smoothing when that
// phase rewrites every interior cell of its colour -- no periodic faces, or
// the wrapped phases over the whole interior -- so that the prolongation can
// leave those cells alone (the phase reads only the other colour); -1
// otherwise.
template <typename T>
int bmg_ctx::postPhaseColour(int l) const
{
    if (opts.n_sweeps_post <= 0 || opts.smoother != BMG_RBGS ||
        fusedLevel<T>(l))
        return -1;
    const int wrap = wrapAxes<T>(l);
    if (anyPeriodic() && !wrap) return -1;
    int cols[4];
    return phase_colours(interior(l, wrap ? 0 : 1), cols) > 0 ? cols[0] : -1;
}

template <typename T>
void bmg_ctx::coarseSolve()
{
    const int l = coarseLevel();
    if (!prepared || preparedCoarse != l)
        throw Error(BMG_ERR_STATE, "coarse level not factorised");
    // the FP32 cycles solve by the inverse; FP64 keeps the reference's
    // substitutions, bit for bit
    if (std::is_same<T, float>::value && dCoarseInv)
        launch_coarse_inv(lv[l].L, op<T>(l), X<T>(l), cst(B<T>(l)), dCoarseInv,
                          nc, l == 0, &bcval, stream);
    else
        launch_coarse_solve(lv[l].L, op<T>(l), X<T>(l), cst(B<T>(l)), dLU,
                            dPiv, nc, l == 0, &bcval, stream);
}

// MG::cycle (MG.C:19-129).  On the coarsest level in direct mode the
// pre-smoothing is skipped: APLU::solve overwrites every interior value and
// reads ghosts only through eliminated (zero) couplings, so the result is the
// same.
// Structural position of a level's launches for batched merges: the part
// before the coarser levels at height h above the coarse level, the coarse
// level, and the part after them; increasing along a V-cycle.
static long cycleTag(int height, bool after)
{
    return after ? 1000 + height : 1000 - height;
}

template <typename T>
void bmg_ctx::cycle(std::vector<int>& visits, int l, bool zeroX)
{
    Level& q = lv[l];
    const int lc = coarseLevel();
    record_tag(cycleTag(lc - l, false));
    if (singular() && l == lc)
        launch_set_cell<T>(B<T>(l), ((long)1*q.L.P[1] + 1)*q.L.Hz + 0,
                           (1 + 1 + 1) & 1, T(0), stream);
    const bool direct = opts.coarse_mode == BMG_COARSE_DIRECT;
    const bool pre = opts.n_sweeps_pre > 0 && !(l == lc && direct);
    bool firstDone = false, firstZero = false;
    if (l > 0 || zeroX)
    {
        // x = 0 and, with red-black smoothing, the first phase from it in
        // one launch, on levels of at most 64^3 cells, where a launch costs
        // more than the work; on larger ones the tiled phase after a plain
        // fill is faster (measured at 256^3 in FP32: 11.6 + 32 us against
        // 62.6 us in one launch, whose linear element order reuses the
        // face coefficients less)
        const long cells = long(q.L.N[0])*q.L.N[1]*q.L.N[2];
        if (pre && std::is_same<T, float>::value && firstPhaseRewrites<T>(l))
        {
            // FP32 cycles: no fill; the first phase takes its neighbours as
            // 0 and the second rewrites every cell of the other colour
            // before anything reads it.  Ghosts behind eliminated faces keep
            // what an earlier cycle left (finite: every solve start zeroes
            // the FP32 fields, clearStaleCorrections) and are multiplied by
            // 0, so results may differ from the filled form in the sign of
            // a zero; no bitwise contract here
            firstZero = true;
        }
        else if (pre && opts.smoother == BMG_RBGS && !fusedLevel<T>(l) &&
                 cells <= (1L << 18))
        {
            const int wrap = wrapAxes<T>(l);
            Region rg = interior(l, wrap ? 0 : 1);
            rg.wrap = wrap;
            int cols[4];
            if (phase_colours(rg, cols) > 0)
            {
                launch_zero_phase<T>(q.L, op<T>(l), X<T>(l), cst(B<T>(l)),
                                     cols[0], rg, stream);
                firstDone = true;
            }
        }
        if (!firstDone && !firstZero)
            launch_fill<T>(X<T>(l), q.L.ncol, T(0), stream);
    }
    if (pre)
        smooth<T>(l, opts.n_sweeps_pre, firstDone, firstZero);
    if (l < lc)
    {
        int reps = 1;
        if (opts.cycle == BMG_WCYCLE) reps = 1 + (l > 0);
        else if (opts.cycle == BMG_FCYCLE) reps = 1 + (l > 0 && visits[l] == 0);
        for (int rep = 0; rep < reps; rep++)
        {
            // the phases of a wrapped level leave its periodic ghosts current
            if (rep > 0 || l > 0)
                residual<T>(l, wrapAxes<T>(l) == 0, false, nullptr);
            restrictTo<T>(l);
            cycle<T>(visits, l + 1, true);
            record_tag(cycleTag(lc - l, true));
            prolongFrom<T>(l, postPhaseColour<T>(l));
            if (opts.n_sweeps_post > 0)
                smooth<T>(l, opts.n_sweeps_post);
        }
    }
    else
    {
        if (direct)
            coarseSolve<T>();
        else
            smooth<T>(l, std::max(opts.n_sweeps_post, 2));
    }
    visits[l]++;
}

// APLU::prepare (APLU.C:26-89): dense matrix of the coarse level from the
// CSR rows of linearSystemAggregation (out-of-system couplings removed,
// repeated columns merged, linearSystemAggregation.C:260-285), then Crout LU
// with implicit partial pivoting (Press et al., ludcmp), the algorithm of
// OpenFOAM's LUDecompose.
void bmg_ctx::prepareCoarse()
{
    const int l = coarseLevel();
    Level& q = lv[l];
    const LevelDims& L = q.L;
    const int n = L.N[0]*L.N[1]*L.N[2];
    if (n > 4096)
        throw Error(BMG_ERR_ARG, "coarse level too large for the direct "
                                 "solver (more than 4096 cells)");
    std::vector<double> h0(L.ncol*7), h1(L.ncol*7);
    HIPCK(hipStreamSynchronize(stream));
    HIPCK(hipMemcpy(h0.data(), q.a7[0], h0.size()*sizeof(double),
                    hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(h1.data(), q.a7[1], h1.size()*sizeof(double),
                    hipMemcpyDeviceToHost));
    auto number = [&](int I, int J, int K) -> long
    {
        int g[3] = {I - 1, J - 1, K - 1};
        for (int d = 0; d < 3; d++)
        {
            if (g[d] < 0 || g[d] >= L.N[d])
            {
                if (!periodic(2*d + (g[d] >= L.N[d]))) return -1;
                g[d] = (g[d] + L.N[d]) % L.N[d];
            }
        }
        return (long(g[0])*L.N[1] + g[1])*L.N[2] + g[2];
    };
    static const int off[7][3] = {{0,0,0},{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},
                                  {0,0,-1},{0,0,1}};
    std::vector<double> D(size_t(n)*n, 0.0);
    for (int I = 1; I <= L.N[0]; I++)
    for (int J = 1; J <= L.N[1]; J++)
    for (int K = 1; K <= L.N[2]; K++)
    {
        const long row = number(I, J, K);
        const int colour = (I + J + K) & 1;
        const long idx = ((long)I*L.P[1] + J)*L.Hz + (K >> 1);
        const std::vector<double>& h = colour ? h1 : h0;
        double a[7];
        long col[7];
        for (int s = 0; s < 7; s++)
        {
            a[s] = h[s*L.ncol + idx];
            col[s] = number(I + off[s][0], J + off[s][1], K + off[s][2]);
            if (col[s] < 0 || col[s] >= n) a[s] = 0;
        }
        for (int s = 1; s < 7; s++)
        {
            int t = 0;
            while (col[t] != col[s]) t++;
            if (t < s) { a[t] += a[s]; a[s] = 0; }
        }
        for (int s = 0; s < 7; s++)
            if (a[s] != 0) D[size_t(row)*n + col[s]] = a[s];
    }
    // Crout LU, implicit pivoting
    std::vector<double> vv(n);
    std::vector<int> piv(n, 0);
    for (int i = 0; i < n; i++)
    {
        double big = 0.0;
        for (int j = 0; j < n; j++) big = std::max(big, std::fabs(D[size_t(i)*n + j]));
        if (big == 0.0)
            throw Error(BMG_ERR_SINGULAR, "coarse matrix has an empty row");
        vv[i] = 1.0/big;
    }
    for (int j = 0; j < n; j++)
    {
        for (int i = 0; i < j; i++)
        {
            double sum = D[size_t(i)*n + j];
            for (int k = 0; k < i; k++) sum -= D[size_t(i)*n + k]*D[size_t(k)*n + j];
            D[size_t(i)*n + j] = sum;
        }
        int imax = j;
        double big = 0.0;
        for (int i = j; i < n; i++)
        {
            double sum = D[size_t(i)*n + j];
            for (int k = 0; k < j; k++) sum -= D[size_t(i)*n + k]*D[size_t(k)*n + j];
            D[size_t(i)*n + j] = sum;
            const double t = vv[i]*std::fabs(sum);
            if (t >= big) { big = t; imax = i; }
        }
        piv[j] = imax;
        if (j != imax)
        {
            for (int k = 0; k < n; k++)
                std::swap(D[size_t(imax)*n + k], D[size_t(j)*n + k]);
            vv[imax] = vv[j];
        }
        if (D[size_t(j)*n + j] == 0.0) D[size_t(j)*n + j] = 1e-15;
        if (j != n - 1)
        {
            const double rr = 1.0/D[size_t(j)*n + j];
            for (int i = j + 1; i < n; i++) D[size_t(i)*n + j] *= rr;
        }
    }
    // the inverse for the FP32 cycles (at most 64 cells): the factors'
    // substitutions (LUBacksubstitute's order) on every unit vector
    devFree(dCoarseInv);
    if (n <= 64)
    {
        std::vector<double> G(64*64, 0.0), e(n);
        for (int k = 0; k < n; k++)
        {
            std::fill(e.begin(), e.end(), 0.0);
            e[k] = 1.0;
            int ii = 0;
            for (int i = 0; i < n; i++)
            {
                const int ip = piv[i];
                double sum = e[ip];
                e[ip] = e[i];
                if (ii)
                    for (int j = ii - 1; j < i; j++) sum -= D[size_t(i)*n + j]*e[j];
                else if (sum != 0.0)
                    ii = i + 1;
                e[i] = sum;
            }
            for (int i = n - 1; i >= 0; i--)
            {
                double sum = e[i];
                for (int j = i + 1; j < n; j++) sum -= D[size_t(i)*n + j]*e[j];
                e[i] = sum/D[size_t(i)*n + i];
            }
            // column k of the inverse; lane i reads entry (i, j) at j*64 + i
            for (int i = 0; i < n; i++) G[size_t(k)*64 + i] = e[i];
        }
        dCoarseInv = devUpload(G);
        gen++;
    }
    // a coarse level that fits one wavefront is solved from padded
    // column-major factors and the final row permutation (the pivot swaps
    // applied in order to the identity)
    const int w = coarse_wave_width(n);
    if (w)
    {
        std::vector<double> F(size_t(w)*w, 0.0);
        for (int i = 0; i < w; i++)
            for (int j = 0; j < w; j++)
                F[size_t(j)*w + i] = i < n && j < n ? D[size_t(i)*n + j]
                                                    : double(i == j);
        std::vector<int> perm(n);
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int i = 0; i < n; i++) std::swap(perm[piv[i]], perm[i]);
        D.swap(F);
        piv.swap(perm);
    }
    if (dLU && nc == n)
    {
        HIPCK(hipMemcpy(dLU, D.data(), D.size()*sizeof(double),
                        hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(dPiv, piv.data(), piv.size()*sizeof(int),
                        hipMemcpyHostToDevice));
    }
    else
    {
        devFree(dLU);
        devFree(dPiv);
        dLU = devUpload(D);
        dPiv = devUpload(piv);
        gen++;
    }
    nc = n;
    prepared = true;
    preparedCoarse = l;
}

// MG.C:431-459: singularity pin of the coarse level, ghost elimination
// (with the level-0 inhomogeneous source terms), coarse factorisation.
void bmg_ctx::prepare(bool constMatrix)
{
    for (auto& q : lv)
        if (!q.haveOp)
            throw Error(BMG_ERR_STATE, "operator not set on every level");
    const int lc = coarseLevel();
    if (lc < 1 && nLevels() > 1)
        throw Error(BMG_ERR_ARG, "coarse level is set too high");
    Level& qc = lv[lc];
    if (qc.opkind != OP_STENCIL)
    {
        // the pinned coarse level needs the stored stencil
        for (int c = 0; c < 2; c++)
        {
            qc.a7[c] = devAlloc<double>(qc.L.ncol*7);
            HIPCK(hipMemsetAsync(qc.a7[c], 0, qc.L.ncol*7*sizeof(double),
                                 stream));
        }
        launch_materialize_stencil(qc.L, op<double>(lc), pr(qc.a7), stream);
        qc.opkind = OP_STENCIL;
        mixedReady = false;
        gen++;
    }
    if (singular())
        launch_pin(qc.L, pr(qc.a7), stream);
    if (!eliminated)
    {
        // the boundary terms of b are folded at solve start (after b is
        // loaded); a stored level-0 stencil keeps its ghost coefficients for
        // them here, before elimination zeroes them
        for (int f = 0; f < 6; f++)
        {
            bcval.g[f] = nullptr;
            if (lv[0].opkind != OP_STENCIL || periodic(f) || !gsave[f])
                continue;
            bmg::launch_save_ghost_coeffs(lv[0].L, cpr(lv[0].a7), f, gsave[f],
                                          stream);
            bcval.g[f] = gsave[f];
        }
        for (auto& q : lv)
            if (q.opkind == OP_STENCIL)
                launch_eliminate(q.L, pr(q.a7), stream);
        eliminated = true;
        mixedReady = false;
    }
    HIPCK(hipGetLastError());
    if (opts.coarse_mode == BMG_COARSE_DIRECT &&
        (!constMatrix || !prepared || preparedCoarse != lc))
        prepareCoarse();
}

// FP32 copies of the operator and transfer data for mixed-precision cycles.
void bmg_ctx::ensureMixed()
{
    if (mixedReady) return;
    gen++;
    for (int l = 0; l < nLevels(); l++)
    {
        Level& q = lv[l];
        const long n = q.L.ncol;
        for (int c = 0; c < 2; c++)
        {
            if (!q.xf[c]) q.xf[c] = devAlloc<float>(n);
            if (!q.bf[c]) q.bf[c] = devAlloc<float>(n);
            if (l > 0 && !q.rf[c]) q.rf[c] = devAlloc<float>(n);
        }
        if (l == 0)
            for (int c = 0; c < 2; c++) q.rf[c] = q.bf[c];  // r aliases b
        for (int d = 0; d < 3; d++)
            if (periodic(2*d))
                for (int k = 0; k < 2; k++)
                    if (!q.snapf[d][k])
                        q.snapf[d][k] = devAlloc<float>(q.planeCells(d));
        launch_fill<float>(pr(q.xf), n, 0.0f, stream);
        launch_fill<float>(pr(q.bf), n, 0.0f, stream);
        if (q.opkind == OP_STENCIL)
        {
            for (int c = 0; c < 2; c++)
                if (!q.a7f[c]) q.a7f[c] = devAlloc<float>(n*7);
            Pair<const double> in{{q.a7[0], q.a7[1]}};
            launch_to_float(n*7, in, pr(q.a7f), stream);
        }
        else if (q.opkind == OP_FACES)
        {
            // FP16 faces scaled so that the largest is 2^14 (the FP16
            // normal range then spans nine decades below it)
            for (int c = 0; c < 2; c++)
                if (!q.wff[c]) q.wff[c] = devAlloc<_Float16>(n*3);
            Pair<const double> in{{q.wf[0], q.wf[1]}};
            const double mx = device_absmax(n*3, in, partials, npartials,
                                            stream);
            const double fs = mx > 0.0 ? mx/16384.0 : 1.0;
            q.wfs = float(fs);
            launch_to_half(n*3, in, 1.0/double(q.wfs), pr(q.wff), stream);
        }
        if (q.opkind != OP_STENCIL && q.dg[0])
        {
            for (int c = 0; c < 2; c++)
                if (!q.dgf[c]) q.dgf[c] = devAlloc<float>(n);
            Pair<const double> in{{q.dg[0], q.dg[1]}};
            launch_to_float(n, in, pr(q.dgf), stream);
        }
        for (int d = 0; d < 3; d++)
        {
            auto f32 = [](const std::vector<double>& v)
            {
                std::vector<float> o(v.size());
                for (size_t i = 0; i < v.size(); i++) o[i] = float(v[i]);
                return o;
            };
            if (!q.fw[d]) q.fw[d] = devUpload(f32(q.ax[d].w));
            if (!q.fdl[d]) q.fdl[d] = devUpload(f32(q.ax[d].dl));
        }
    }
    // transfer u arrays in FP32 (host copies were uploaded in FP64)
    for (int l = 0; l < nLevels(); l++)
    {
        Level& q = lv[l];
        for (int d = 0; d < 3; d++)
        {
            auto conv = [&](double* src, float*& dst, int n)
            {
                if (!src || dst) return;
                std::vector<double> h(n);
                HIPCK(hipMemcpy(h.data(), src, n*sizeof(double),
                                hipMemcpyDeviceToHost));
                std::vector<float> f(n);
                for (int i = 0; i < n; i++) f[i] = float(h[i]);
                dst = devUpload(f);
            };
            conv(q.ru[d], q.ruf[d], q.L.P[d]);
            conv(q.pu[d], q.puf[d], q.L.P[d]);
        }
    }
    HIPCK(hipGetLastError());
    mixedReady = true;
}

// One MG iteration (MG.C:207-230): the cycle, the level-0 residual with its
// norm, and the copy of the norm to pinned host memory.  No synchronisation
// inside, so the sequence can be captured into a HIP graph.
void bmg_ctx::enqueueIteration(bool mixed)
{
    Level& q0 = lv[0];
    std::vector<int> visits(nLevels(), 0);
    record_tag(cycleTag(coarseLevel(), false));
    if (!mixed)
        cycle<double>(visits, 0);
    else
    {
        // defect equation A e = r in FP32, e0 = 0; x += e in FP64
        Pair<const double> r64{{q0.r[0], q0.r[1]}};
        launch_to_float(q0.L.ncol, r64, pr(q0.bf), stream);
        cycle<float>(visits, 0, true);
        launch_add_correction(q0.L, pr(q0.x), cpr(q0.xf), stream);
    }
    int nb = 0;
    residual<double>(0, true, true, &nb);
    launch_reduce_final(partials, nb, dScalar, stream, hScalarDev);
}

// Runs one iteration and returns the sum of squared level-0 residuals.
double bmg_ctx::iterate(bool mixed)
{
    if (useGraphs)
    {
        ensureSmootherBuffers(mixed);
        if (!graphExec || graphGen != gen || graphMixed != int(mixed))
        {
            if (graphExec) { HIPCK(hipGraphExecDestroy(graphExec)); graphExec = nullptr; }
            if (graph) { HIPCK(hipGraphDestroy(graph)); graph = nullptr; }
            HIPCK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try
            {
                enqueueIteration(mixed);
            }
            catch (...)
            {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCK(hipStreamEndCapture(stream, &graph));
            HIPCK(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0));
            graphGen = gen;
            graphMixed = int(mixed);
        }
        HIPCK(hipGraphLaunch(graphExec, stream));
    }
    else
        enqueueIteration(mixed);
    HIPCK(hipStreamSynchronize(stream));
    return *hScalar;
}

// z = M(r): one MG cycle on A z = r with z0 = 0, in FP64 or in FP32.  The
// caller has staged r as the level-0 right-hand side of the cycle (and, in
// FP64, as the residual its first repetition restricts, MG.C:72-77); z is
// left in the level-0 x (FP64) or xf (FP32).
void bmg_ctx::enqueuePrecondition(bool mixed)
{
    std::vector<int> visits(nLevels(), 0);
    if (!mixed)
        cycle<double>(visits, 0, true);
    else
        cycle<float>(visits, 0, true);
}

// One flexible-CG iteration with its scalars on the device (sc = dScalar):
// q = A p with p.q, p.r; r -= alpha q with r.r and r staged for the
// preconditioner; z = M r; z.q; x += alpha p and p = z + beta p in one pass
// (x is complete again when the iteration ends); the four scalars copied to
// the pinned host buffer.
void bmg_ctx::enqueueFCGIteration(int mode)
{
    Level& q0 = lv[0];
    const OpView<double> o = op<double>(0);
    Pair<double> X{{kx[0], kx[1]}}, Rv{{kr[0], kr[1]}}, Pv{{kp[0], kp[1]}},
        Qv{{kq[0], kq[1]}};
    // batched merges: the part before the cycle, the cycle's own positions,
    // the part after it
    record_tag(0);
    if (anyPeriodic())
        launch_bc_fill(q0.L, o, Pv, 0, 1, &bcval, stream);
    const int nb = launch_apply_dots(q0.L, o, cst(Pv), Qv, cst(Rv), partials,
                                     stream);
    launch_reduce_rows(partials, nb, 2, dScalar, stream, hScalarDev);
    const Pair<float> rf = mode == 1 ? pr(q0.bf)
                                     : Pair<float>{{nullptr, nullptr}};
    const Pair<double> r0 = mode == 0 ? pr(q0.r)
                                      : Pair<double>{{nullptr, nullptr}};
    const int nbu = launch_fcg_update(q0.L, dScalar, Rv, cst(Qv), rf, r0,
                                      pr(q0.b), partials, stream);
    // r.r is needed on the device only by the learned preconditioner; else
    // its partials wait (the cycle takes no norms) and are summed together
    // with those of z.q, one launch
    if (mode == 2)
    {
        launch_reduce_rows(partials, nbu, 1, dScalar + 2, stream,
                           hScalarDev + 2);
        enqueueLearned(dScalar + 2, 0.0, pr(q0.x));
    }
    else
        enqueuePrecondition(mode == 1);
    record_tag(3000);
    double* const zqPart = mode == 2 ? partials : partials + nbu;
    const int nbz =
        mode == 1 ? launch_dot_zq<float>(q0.L, cpr(q0.xf), cst(Qv), zqPart,
                                         stream)
                  : launch_dot_zq<double>(q0.L, cpr(q0.x), cst(Qv), zqPart,
                                          stream);
    if (mode == 2)
        launch_reduce_rows(partials, nbz, 1, dScalar + 3, stream,
                           hScalarDev + 3);
    else
    {
        if (nbz != nbu) throw Error(BMG_ERR_STATE, "FCG partial rows differ");
        launch_reduce_rows(partials, nbu, 2, dScalar + 2, stream,
                           hScalarDev + 2);
    }
    if (mode == 1)
        launch_update_p<float>(q0.L, dScalar, cpr(q0.xf), Pv, X, stream);
    else
        launch_update_p<double>(q0.L, dScalar, cpr(q0.x), Pv, X, stream);
    if (mode != 2 && deflating()) enqueueDeflate();
    HIPCK(hipGetLastError());
}

// z = U(r / rms(r)) * rms(r) / h on the level-0 interior (the learned
// preconditioner, bmg_unet.h); rms from *ssDev or ssHost (sum of squares).
void bmg_ctx::enqueueLearned(const double* ssDev, double ssHost,
                             Pair<double> z)
{
    Level& q0 = lv[0];
    // z = U * rms / (h s), h the cell width along the first axis and s the
    // geometric-mean coefficient
    const double len = lv[0].ax[0].v.back() - lv[0].ax[0].v.front();
    const double invh = double(q0.L.N[0])/(len*coefScale);
    launch_split_to_dense(q0.L, Pair<const double>{{kr[0], kr[1]}}, ssDev,
                          ssHost, uIn, stream);
    bmg::unet::net_forward(unet, q0.L.N[0], q0.L.N[1], q0.L.N[2], uIn, uCoef,
                           uOut, stream);
    launch_dense_to_split(q0.L, uOut, ssDev, ssHost, invh, z, stream);
}

// Buffers and coefficient channel of the learned preconditioner (outside
// any graph capture: this allocates).
void bmg_ctx::prepareLearned()
{
    if (!unet)
        throw Error(BMG_ERR_STATE, "learned Krylov mode without a network "
                                   "(bmg_set_learned)");
    Level& q0 = lv[0];
    for (int d = 0; d < 3; d++)
        if (q0.L.N[d] % 8)
            throw Error(BMG_ERR_ARG, "the learned preconditioner needs level-0 "
                                     "extents that are multiples of 8");
    const size_t M = size_t(q0.cells());
    if (!uIn)
    {
        uIn = devAlloc<float>(M);
        uCoef = devAlloc<float>(M);
        uOut = devAlloc<float>(M);
        coefUploaded = false;
        gen++;
    }
    if (!coefUploaded)
    {
        if (coefChannel.size() != M) coefChannel.assign(M, 1.0f);
        HIPCK(hipMemcpy(uCoef, coefChannel.data(), M*sizeof(float),
                        hipMemcpyHostToDevice));
        coefUploaded = true;
    }
    try
    {
        bmg::unet::net_prepare(unet, q0.L.N[0], q0.L.N[1], q0.L.N[2]);
    }
    catch (const std::exception& e)
    {
        throw Error(BMG_ERR_HIP, e.what());
    }
}

void bmg_ctx::fcgIteration(int mode)
{
    if (!useGraphs)
    {
        enqueueFCGIteration(mode);
        HIPCK(hipStreamSynchronize(stream));
        return;
    }
    if (mode != 2) ensureSmootherBuffers(mode == 1);
    if (!fgraphExec || fgraphGen != gen || fgraphMode != mode)
    {
        if (fgraphExec) { HIPCK(hipGraphExecDestroy(fgraphExec)); fgraphExec = nullptr; }
        if (fgraph) { HIPCK(hipGraphDestroy(fgraph)); fgraph = nullptr; }
        HIPCK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try
        {
            enqueueFCGIteration(mode);
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCK(hipStreamEndCapture(stream, &fgraph));
        HIPCK(hipGraphInstantiate(&fgraphExec, fgraph, nullptr, nullptr, 0));
        fgraphGen = gen;
        fgraphMode = mode;
    }
    HIPCK(hipGraphLaunch(fgraphExec, stream));
    HIPCK(hipStreamSynchronize(stream));
}

void bmg_ctx::precondition(bool mixed)
{
    if (!useGraphs)
    {
        enqueuePrecondition(mixed);
        return;
    }
    ensureSmootherBuffers(mixed);
    if (!pgraphExec || pgraphGen != gen || pgraphMixed != int(mixed))
    {
        if (pgraphExec) { HIPCK(hipGraphExecDestroy(pgraphExec)); pgraphExec = nullptr; }
        if (pgraph) { HIPCK(hipGraphDestroy(pgraph)); pgraph = nullptr; }
        HIPCK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try
        {
            enqueuePrecondition(mixed);
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCK(hipStreamEndCapture(stream, &pgraph));
        HIPCK(hipGraphInstantiate(&pgraphExec, pgraph, nullptr, nullptr, 0));
        pgraphGen = gen;
        pgraphMixed = int(mixed);
    }
    HIPCK(hipGraphLaunch(pgraphExec, stream));
}

// ---------------------------------------------------------------------------
// Deflation of flexible CG by indicator vectors of high-coefficient cell
// sets (opts.deflation).  Labels are found on the host from the face lambda
// arrays when the operator is assembled: a cell's coefficient is the mean of
// its six face lambdas; with a contrast of at least 100 between the extremes
// the cells above their geometric mean are grouped into face-connected sets
// (periodic faces wrap), and the largest sets of at least 8 cells, at most
// 64, are kept.
static const int DEFL_MAX = 64;
static const long DEFL_MIN_CELLS = 8;

void bmg_ctx::deflationLabels(const double* lx, const double* ly,
                              const double* lz)
{
    deflLabel.clear();
    deflM = 0;
    const LevelDims& L0 = lv[0].L;
    const int N0 = L0.N[0], N1 = L0.N[1], N2 = L0.N[2];
    const long P1 = L0.P[1], P2 = L0.P[2];
    const long n = long(N0)*N1*N2;
    std::vector<double> m(n);
    double lo = 1e300, hi = 0.0;
    for (int i = 1; i <= N0; i++)
        for (int j = 1; j <= N1; j++)
            for (int k = 1; k <= N2; k++)
            {
                const long p = (i*P1 + j)*P2 + k;
                const double v = (lx[p] + lx[p + P1*P2] + ly[p] + ly[p + P2] +
                                  lz[p] + lz[p + 1])/6.0;
                m[(long(i - 1)*N1 + (j - 1))*N2 + (k - 1)] = v;
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
    if (!(lo > 0.0) || !(hi >= 100.0*lo)) return;
    const double theta = std::sqrt(lo*hi);
    std::vector<int> lab(n, -1);
    std::vector<long> size, stack;
    const int N[3] = {N0, N1, N2};
    for (long c = 0; c < n; c++)
    {
        if (lab[c] != -1 || !(m[c] > theta)) continue;
        const int id = int(size.size());
        long cnt = 0;
        lab[c] = id;
        stack.assign(1, c);
        while (!stack.empty())
        {
            const long u = stack.back();
            stack.pop_back();
            cnt++;
            const int g[3] = {int(u/(long(N1)*N2)), int((u/N2) % N1),
                              int(u % N2)};
            for (int d = 0; d < 3; d++)
                for (int sgn = -1; sgn <= 1; sgn += 2)
                {
                    int q[3] = {g[0], g[1], g[2]};
                    q[d] += sgn;
                    if (q[d] < 0 || q[d] >= N[d])
                    {
                        if (!periodic(2*d + (q[d] >= N[d]))) continue;
                        q[d] = (q[d] + N[d]) % N[d];
                    }
                    const long v = (long(q[0])*N1 + q[1])*N2 + q[2];
                    if (lab[v] == -1 && m[v] > theta)
                    {
                        lab[v] = id;
                        stack.push_back(v);
                    }
                }
        }
        size.push_back(cnt);
    }
    // keep the largest sets of at least DEFL_MIN_CELLS cells
    std::vector<int> order(size.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = int(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return size[a] > size[b]; });
    std::vector<int> remap(size.size(), -1);
    int kept = 0;
    for (int id : order)
    {
        if (kept >= DEFL_MAX || size[id] < DEFL_MIN_CELLS) break;
        remap[id] = kept++;
    }
    if (kept == 0) return;
    for (long c = 0; c < n; c++)
        if (lab[c] >= 0) lab[c] = remap[lab[c]];
    deflLabel.swap(lab);
    deflM = kept;
}

bool bmg_ctx::deflating() const
{
    return opts.deflation != 0 && deflM > 0 && deflOpGen == opGen &&
           dDeflEinv != nullptr && launch_recorder() == nullptr;
}

// Device data of the current operator's deflation: the entries (A z_k)_i
// (counted, written, sorted by component and cell on the host), E = Z^T A Z
// from them, its inverse (Gauss-Jordan with partial pivoting), and the cell
// lists of the components.
void bmg_ctx::ensureDeflation()
{
    if (opts.deflation == 0 || deflM == 0 || deflOpGen == opGen) return;
    freeDeflation();
    Level& q0 = lv[0];
    const LevelDims& L = q0.L;
    const long n = long(L.N[0])*L.N[1]*L.N[2];
    const int m = deflM;
    int* dLab = devAlloc<int>(n);
    unsigned* dCount = devAlloc<unsigned>(1);
    std::vector<unsigned> hc;
    std::vector<int> hk;
    std::vector<double> hv;
    try
    {
        HIPCK(hipMemcpy(dLab, deflLabel.data(), n*sizeof(int),
                        hipMemcpyHostToDevice));
        HIPCK(hipMemsetAsync(dCount, 0, sizeof(unsigned), stream));
        launch_defl_entries(L, op<double>(0), dLab, dCount, 0, nullptr,
                            nullptr, nullptr, stream);
        unsigned cnt = 0;
        HIPCK(hipMemcpyAsync(&cnt, dCount, sizeof(unsigned),
                             hipMemcpyDeviceToHost, stream));
        HIPCK(hipStreamSynchronize(stream));
        unsigned* dc = devAlloc<unsigned>(std::max(cnt, 1u));
        int* dk = devAlloc<int>(std::max(cnt, 1u));
        double* dv = devAlloc<double>(std::max(cnt, 1u));
        HIPCK(hipMemsetAsync(dCount, 0, sizeof(unsigned), stream));
        launch_defl_entries(L, op<double>(0), dLab, dCount, cnt, dc, dk, dv,
                            stream);
        hc.resize(cnt);
        hk.resize(cnt);
        hv.resize(cnt);
        HIPCK(hipMemcpyAsync(hc.data(), dc, cnt*sizeof(unsigned),
                             hipMemcpyDeviceToHost, stream));
        HIPCK(hipMemcpyAsync(hk.data(), dk, cnt*sizeof(int),
                             hipMemcpyDeviceToHost, stream));
        HIPCK(hipMemcpyAsync(hv.data(), dv, cnt*sizeof(double),
                             hipMemcpyDeviceToHost, stream));
        HIPCK(hipStreamSynchronize(stream));
        devFree(dc);
        devFree(dk);
        devFree(dv);
    }
    catch (...)
    {
        devFree(dLab);
        devFree(dCount);
        throw;
    }
    devFree(dLab);
    devFree(dCount);
    // entries by component, then cell
    std::vector<size_t> ord(hc.size());
    for (size_t e = 0; e < ord.size(); e++) ord[e] = e;
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        return hk[a] != hk[b] ? hk[a] < hk[b] : hc[a] < hc[b];
    });
    const long Hz = L.Hz, P1 = L.P[1];
    auto labelOf = [&](unsigned code) -> int {
        const long idx = long(code & 0x7fffffffu);
        const int colour = int(code >> 31);
        const int I = int(idx/(P1*Hz));
        const long rem = idx - I*P1*Hz;
        const int J = int(rem/Hz);
        const int p = int(rem - J*Hz);
        const int K = 2*p + ((colour - I - J) & 1);
        return deflLabel[(long(I - 1)*L.N[1] + (J - 1))*L.N[2] + (K - 1)];
    };
    std::vector<unsigned> ec(ord.size());
    std::vector<double> ev(ord.size());
    std::vector<int> eoff(m + 1, 0);   // entries of component k
    std::vector<double> E(size_t(m)*m, 0.0);
    for (size_t e = 0; e < ord.size(); e++)
    {
        const size_t o = ord[e];
        ec[e] = hc[o];
        ev[e] = hv[o];
        eoff[hk[o] + 1]++;
        const int j = labelOf(hc[o]);
        if (j >= 0) E[size_t(j)*m + hk[o]] += hv[o];
    }
    for (int k = 0; k < m; k++) eoff[k + 1] += eoff[k];
    // E^-1 by Gauss-Jordan with partial pivoting
    std::vector<double> A(E), Inv(size_t(m)*m, 0.0);
    for (int i = 0; i < m; i++) Inv[size_t(i)*m + i] = 1.0;
    for (int col = 0; col < m; col++)
    {
        int piv = col;
        for (int r = col + 1; r < m; r++)
            if (std::fabs(A[size_t(r)*m + col]) > std::fabs(A[size_t(piv)*m + col]))
                piv = r;
        if (!(std::fabs(A[size_t(piv)*m + col]) > 0.0))
        {
            deflM = 0;   // singular: solve without deflation
            return;
        }
        if (piv != col)
            for (int c2 = 0; c2 < m; c2++)
            {
                std::swap(A[size_t(piv)*m + c2], A[size_t(col)*m + c2]);
                std::swap(Inv[size_t(piv)*m + c2], Inv[size_t(col)*m + c2]);
            }
        const double d = A[size_t(col)*m + col];
        for (int c2 = 0; c2 < m; c2++)
        {
            A[size_t(col)*m + c2] /= d;
            Inv[size_t(col)*m + c2] /= d;
        }
        for (int r = 0; r < m; r++)
        {
            if (r == col) continue;
            const double f = A[size_t(r)*m + col];
            if (f == 0.0) continue;
            for (int c2 = 0; c2 < m; c2++)
            {
                A[size_t(r)*m + c2] -= f*A[size_t(col)*m + c2];
                Inv[size_t(r)*m + c2] -= f*Inv[size_t(col)*m + c2];
            }
        }
    }
    // component cells
    std::vector<std::vector<unsigned>> cells(m);
    for (int I = 1; I <= L.N[0]; I++)
        for (int J = 1; J <= L.N[1]; J++)
            for (int K = 1; K <= L.N[2]; K++)
            {
                const int k = deflLabel[(long(I - 1)*L.N[1] + (J - 1))*L.N[2] +
                                        (K - 1)];
                if (k < 0) continue;
                const unsigned idx = unsigned((I*P1 + J)*Hz + (K >> 1));
                cells[k].push_back((unsigned((I + J + K) & 1) << 31) | idx);
            }
    std::vector<unsigned> cc;
    std::vector<int> ccomp, coff(m + 1, 0);
    for (int k = 0; k < m; k++)
    {
        cc.insert(cc.end(), cells[k].begin(), cells[k].end());
        ccomp.insert(ccomp.end(), cells[k].size(), k);
        coff[k + 1] = int(cc.size());
    }
    deflNCells = long(cc.size());
    // chunks of at most DEFL_CHUNK list elements within each component
    const int DEFL_CHUNK = 2048;
    auto chunks = [&](const std::vector<int>& off, std::vector<int>& beg,
                      std::vector<int>& end, std::vector<int>& choff) {
        choff.assign(m + 1, 0);
        for (int k = 0; k < m; k++)
        {
            for (int e = off[k]; e < off[k + 1]; e += DEFL_CHUNK)
            {
                beg.push_back(e);
                end.push_back(std::min(off[k + 1], e + DEFL_CHUNK));
            }
            choff[k + 1] = int(beg.size());
        }
    };
    std::vector<int> cb, ce, cco, eb, ee, eco;
    chunks(coff, cb, ce, cco);
    chunks(eoff, eb, ee, eco);
    deflCellChunks = int(cb.size());
    deflEntChunks = int(eb.size());
    auto up = [&](auto*& dst, const auto& v) {
        using E0 = std::remove_reference_t<decltype(*dst)>;
        dst = devAlloc<E0>(std::max<size_t>(v.size(), 1));
        if (!v.empty())
            HIPCK(hipMemcpy(dst, v.data(), v.size()*sizeof(E0),
                            hipMemcpyHostToDevice));
    };
    up(dDeflCell, cc);
    up(dDeflCellComp, ccomp);
    up(dDeflEntCell, ec);
    up(dDeflEntVal, ev);
    up(dDeflCellBeg, cb);
    up(dDeflCellEnd, ce);
    up(dDeflCellChunkOff, cco);
    up(dDeflEntBeg, eb);
    up(dDeflEntEnd, ee);
    up(dDeflEntChunkOff, eco);
    up(dDeflEinv, Inv);
    dDeflPart = devAlloc<double>(size_t(std::max(std::max(deflCellChunks,
                                                          deflEntChunks), 1)));
    deflOpGen = opGen;
    gen++;   // captured graphs must see the new buffers
}

// p -= Z E^-1 (A Z)^T p on the new search direction, in FP64.  With
// P^T = I - Z E^-1 (A Z)^T a projection and every earlier direction already
// A-orthogonal to Z, P^T (z + beta p) = P^T z + beta p and z.q is unchanged,
// so this is deflating the preconditioned residual z, done on the FP64
// direction (z of the mixed cycle is FP32, too coarse for the projection).
void bmg_ctx::enqueueDeflate()
{
    Pair<double> Pv{{kp[0], kp[1]}};
    launch_defl_dot<double>(dDeflEntCell, dDeflEntVal, dDeflEntBeg,
                            dDeflEntEnd, deflEntChunks, cst(Pv), dDeflPart,
                            stream);
    launch_defl_correct<double>(dDeflCell, dDeflCellComp, deflNCells,
                                dDeflEinv, dDeflPart, dDeflEntChunkOff, deflM,
                                -1.0, Pv, stream);
}

// x += Z E^-1 Z^T r on the FCG state: the deflated start.
void bmg_ctx::deflateStart()
{
    Pair<double> X{{kx[0], kx[1]}};
    Pair<const double> Rc{{kr[0], kr[1]}};
    launch_defl_dot<double>(dDeflCell, nullptr, dDeflCellBeg, dDeflCellEnd,
                            deflCellChunks, Rc, dDeflPart, stream);
    launch_defl_correct<double>(dDeflCell, dDeflCellComp, deflNCells, dDeflEinv,
                                dDeflPart, dDeflCellChunkOff, deflM, 1.0, X,
                                stream);
}

// Flexible conjugate gradients (Notay 2000, one stored direction) with one
// MG cycle as the variable preconditioner; the stopping test is the
// reference's (normFactors/checkConvergence, solver.C:13-95) on the updated
// residual, confirmed on the true residual b - A x before returning.  Each
// iteration is one graph launch and one host synchronisation.
void bmg_ctx::runFCG(bool constMatrix, bmg_stats* st)
{
    prepare(constMatrix);
    const bool mixed = opts.precision == BMG_MIXED;
    if (mixed) ensureMixed();
    const bool learned = opts.krylov == BMG_KRYLOV_FCG_LEARNED;
    if (learned) prepareLearned();
    Level& q0 = lv[0];
    const long n = q0.L.ncol;
    prepareFCG();
    if (!learned) ensureDeflation();
    const OpView<double> o = op<double>(0);
    Pair<double> X{{kx[0], kx[1]}}, Bv{{kb[0], kb[1]}}, Rv{{kr[0], kr[1]}},
        Pv{{kp[0], kp[1]}};
    int nb = 0;
    // initial residual without boundary update (MG.C:163) and normFactors
    enqueueFCGStart();
    HIPCK(hipStreamSynchronize(stream));
    double ss = hScalar[0];
    const double cells = double(q0.cells());
    enqueueFCGNormTerms(hScalar[1]/cells);
    HIPCK(hipStreamSynchronize(stream));
    const double nf = std::max(std::sqrt(hScalar[2]), cells*1e-12);
    const double init = std::sqrt(ss)/nf;
    if (!std::isfinite(init))
        throw Error(BMG_ERR_NUMERIC, "initial residual is not finite");
    auto check = [&](double cc) -> int
    {
        return (init > 0.0 && cc > opts.tolerance &&
                cc/(init + 1e-20) > opts.rel_tol) ? 0 : 1;
    };
    double cur = init;
    int conv = check(cur);
    int iter = 0;
    bool restart = true;
    // deflated start x + Z E^-1 Z^T r (the stopping test keeps init)
    if (!learned && deflating() && conv == 0)
    {
        deflateStart();
        if (anyPeriodic())
            launch_bc_fill(q0.L, o, X, 0, 1, &bcval, stream);
        launch_residual(q0.L, o, cst(X), cst(Bv), Rv, partials, &nb, stream);
        ss = reduceFinal(nb);
        cur = std::sqrt(ss)/nf;
        if (!std::isfinite(cur))
            throw Error(BMG_ERR_NUMERIC, "residual norm is not finite");
        conv = check(cur);
    }
    // learned mode: the network's initial guess, kept only if it lowers the
    // residual (x is saved in q meanwhile)
    bool learnedActive = learned;
    int learnedIters = 0, fallbackAt = -1;
    double guessResidual = 0.0;
    if (learned && conv == 0)
    {
        Pair<double> Qv{{kq[0], kq[1]}};
        launch_copy<double>(n, cst(X), Qv, stream);
        enqueueLearned(nullptr, ss, pr(q0.x));
        launch_add_interior(q0.L, X, cpr(q0.x), stream);
        if (anyPeriodic())
            launch_bc_fill(q0.L, o, X, 0, 1, &bcval, stream);
        launch_residual(q0.L, o, cst(X), cst(Bv), Rv, partials, &nb, stream);
        const double ssg = reduceFinal(nb);
        guessResidual = std::sqrt(ssg)/nf;
        if (std::isfinite(ssg) && ssg < ss)
        {
            ss = ssg;
            cur = guessResidual;
            conv = check(cur);
        }
        else
        {
            launch_copy<double>(n, cst(Qv), X, stream);
            launch_residual(q0.L, o, cst(X), cst(Bv), Rv, partials, &nb,
                            stream);
            ss = reduceFinal(nb);
            learnedActive = false;
            fallbackAt = 0;
        }
    }
    while ((iter < opts.max_iter && conv == 0) || iter < opts.min_iter)
    {
        const int mode = learnedActive ? 2 : (mixed ? 1 : 0);
        if (restart)
        {
            // p = z = M r for the current residual
            if (mode == 2)
            {
                enqueueLearned(nullptr, ss, pr(q0.x));
                launch_copy<double>(n, cpr(q0.x), Pv, stream);
            }
            else
            {
                if (useGraphs)
                {
                    // the cycle from its own captured graph
                    stageRestart(mixed);
                    precondition(mixed);
                    finishRestart(mixed);
                    if (deflating()) enqueueDeflate();
                }
                else
                    enqueueFCGRestart(mixed);
            }
            restart = false;
        }
        const double prev = cur;
        fcgIteration(mode);
        const double pq = hScalar[0];
        if (!(pq != 0.0) || !std::isfinite(pq))
            throw Error(BMG_ERR_NUMERIC, "flexible CG breakdown (p.Ap = 0)");
        ss = hScalar[2];
        cur = std::sqrt(ss)/nf;
        if (!std::isfinite(cur))
            throw Error(BMG_ERR_NUMERIC, "residual norm is not finite");
        iter++;
        if (mode == 2)
        {
            learnedIters++;
            // fall back to multigrid cycles once an iteration's residual
            // drop is worse than 0.3
            if (cur > 0.3*prev)
            {
                learnedActive = false;
                fallbackAt = iter;
                restart = true;
            }
        }
        conv = check(cur);
        if (conv && iter >= opts.min_iter)
        {
            // confirm on the true residual (with the boundary update)
            enqueueFCGConfirm();
            HIPCK(hipStreamSynchronize(stream));
            ss = hScalar[0];
            cur = std::sqrt(ss)/nf;
            conv = check(cur);
            if (conv) break;
            restart = true;
        }
    }
    enqueueFCGEnd();
    HIPCK(hipGetLastError());
    if (st)
    {
        st->iters = iter;
        st->initial_residual = init;
        st->final_residual = cur;
        st->learned_iters = learnedIters;
        st->learned_fallback = fallbackAt;
        st->learned_guess_residual = guessResidual;
    }
}

// Flexible-CG buffers (FP64 x, b, r, p, q on level 0), allocated once.
void bmg_ctx::prepareFCG()
{
    if (kx[0]) return;
    const long n = lv[0].L.ncol;
    for (int c = 0; c < 2; c++)
    {
        kx[c] = devAlloc<double>(n); kb[c] = devAlloc<double>(n);
        kr[c] = devAlloc<double>(n); kp[c] = devAlloc<double>(n);
        kq[c] = devAlloc<double>(n);
    }
    gen++;
    for (double** f : {kr, kp, kq})
        launch_fill<double>(Pair<double>{{f[0], f[1]}}, n, 0.0, stream);
}

// FCG start: copies of x and b, r = b - A x without boundary update
// (MG.C:163) with its squared norm (host slot 0) and the sum of x (slot 1).
void bmg_ctx::enqueueFCGStart()
{
    enqueueBoundaryTerms();
    clearStaleCorrections();
    Level& q0 = lv[0];
    const long n = q0.L.ncol;
    Pair<double> X{{kx[0], kx[1]}}, Bv{{kb[0], kb[1]}}, Rv{{kr[0], kr[1]}};
    launch_copy<double>(n, cpr(q0.x), X, stream);
    launch_copy<double>(n, cpr(q0.b), Bv, stream);
    int nb = 0;
    launch_residual(q0.L, op<double>(0), cst(X), cst(Bv), Rv, partials, &nb,
                    stream);
    launch_reduce_final(partials, nb, dScalar, stream, hScalarDev);
    launch_sum<double>(q0.L, cst(X), partials, &nb, stream);
    launch_reduce_final(partials, nb, dScalar + 1, stream, hScalarDev + 1);
}

// normFactors terms of the FCG residual with y = gAverage(x), slot 2.
void bmg_ctx::enqueueFCGNormTerms(double y)
{
    Level& q0 = lv[0];
    Pair<double> Bv{{kb[0], kb[1]}}, Rv{{kr[0], kr[1]}};
    int nb = 0;
    launch_norm_factor_terms<double>(q0.L, op<double>(0), cst(Rv), cst(Bv), y,
                                     partials, &nb, stream);
    launch_reduce_final(partials, nb, dScalar + 2, stream, hScalarDev + 2);
}

// r staged as the cycle's right-hand side (FP32, or FP64 as both the
// right-hand side and the residual the first repetition restricts).
void bmg_ctx::stageRestart(bool mixed)
{
    Level& q0 = lv[0];
    const long n = q0.L.ncol;
    Pair<const double> Rv{{kr[0], kr[1]}};
    if (mixed)
        launch_to_float(n, Rv, pr(q0.bf), stream);
    else
    {
        launch_copy<double>(n, Rv, pr(q0.b), stream);
        launch_copy<double>(n, Rv, pr(q0.r), stream);
    }
}

// p = z, the cycle's result.
void bmg_ctx::finishRestart(bool mixed)
{
    Level& q0 = lv[0];
    const long n = q0.L.ncol;
    Pair<double> Pv{{kp[0], kp[1]}};
    if (mixed)
        launch_to_double(n, Pair<const float>{{q0.xf[0], q0.xf[1]}}, Pv,
                         stream);
    else
        launch_copy<double>(n, cpr(q0.x), Pv, stream);
}

// FCG restart: p = z = M r for the current residual.
void bmg_ctx::enqueueFCGRestart(bool mixed)
{
    record_tag(0);
    stageRestart(mixed);
    enqueuePrecondition(mixed);
    record_tag(3000);
    finishRestart(mixed);
    if (deflating()) enqueueDeflate();
}

// The true residual b - A x with the boundary update, slot 0.
void bmg_ctx::enqueueFCGConfirm()
{
    Level& q0 = lv[0];
    const OpView<double> o = op<double>(0);
    Pair<double> X{{kx[0], kx[1]}}, Bv{{kb[0], kb[1]}}, Rv{{kr[0], kr[1]}};
    if (anyPeriodic())
        launch_bc_fill(q0.L, o, X, 0, 1, &bcval, stream);
    int nb = 0;
    launch_residual(q0.L, o, cst(X), cst(Bv), Rv, partials, &nb, stream);
    launch_reduce_final(partials, nb, dScalar, stream, hScalarDev);
}

// x and b back to the solver's level 0, eliminated ghosts (MG.C:475).
void bmg_ctx::enqueueFCGEnd()
{
    Level& q0 = lv[0];
    const long n = q0.L.ncol;
    Pair<const double> X{{kx[0], kx[1]}}, Bv{{kb[0], kb[1]}};
    launch_copy<double>(n, X, pr(q0.x), stream);
    launch_copy<double>(n, Bv, pr(q0.b), stream);
    launch_bc_fill(q0.L, op<double>(0), pr(q0.x), 2, 1, &bcval, stream);
}

double bmg_ctx::normFactorOf(Pair<const double> x, Pair<const double> r,
                             Pair<const double> b)
{
    Level& q = lv[0];
    const double n = double(q.cells());
    int nb = 0;
    launch_sum<double>(q.L, x, partials, &nb, stream);
    const double sum = reduceFinal(nb);
    const double y = sum/n;
    launch_norm_factor_terms<double>(q.L, op<double>(0), r, b, y, partials,
                                     &nb, stream);
    double f = reduceFinal(nb);
    f = std::sqrt(f);
    return std::max(f, n*1e-12);
}

// Whether prepare() does more than device work this solve: the first
// solve (materialising, eliminating) or a new coarse factorisation.
bool bmg_ctx::prepareNeedsHost(bool constMatrix) const
{
    for (auto& q : lv)
        if (!q.haveOp) return true;
    const int lc = coarseLevel();
    if (lc < 1 && nLevels() > 1) return true;
    if (lv[lc].opkind != OP_STENCIL || !eliminated) return true;
    return opts.coarse_mode == BMG_COARSE_DIRECT &&
           (!constMatrix || !prepared || preparedCoarse != lc);
}

// MG.C:166-172: r = b - A x without BC correction and its squared norm
// (host slot 0), and the sum of x for gAverage (slot 1, solver.C:24).
void bmg_ctx::enqueueBoundaryTerms()
{
    if (foldPending && inhomogeneous())
        launch_bc_source(lv[0].L, op<double>(0), pr(lv[0].b), &bcval, stream);
    foldPending = false;
}

// The FP32 correction fields of every level zeroed, ghosts included.  The
// FP32 cycles leave the ghosts behind eliminated faces as the previous cycle
// left them (they are read only through zero couplings), so values that are
// not finite -- a diverged solve, timing runs on uninitialised fields --
// would otherwise survive into the next solve: 0 times NaN is NaN.
void bmg_ctx::clearStaleCorrections()
{
    if (!mixedReady) return;
    for (int l = 0; l < nLevels(); l++)
        launch_fill<float>(pr(lv[l].xf), lv[l].L.ncol, 0.0f, stream);
}

void bmg_ctx::enqueueSolveStart()
{
    enqueueBoundaryTerms();
    clearStaleCorrections();
    Level& q0 = lv[0];
    launch_fill<double>(pr(q0.r), q0.L.ncol, 0.0, stream);
    int nb = 0;
    launch_residual(q0.L, op<double>(0), cpr(q0.x), cpr(q0.b), pr(q0.r),
                    partials, &nb, stream);
    launch_reduce_final(partials, nb, dScalar, stream, hScalarDev);
    launch_sum<double>(q0.L, cpr(q0.x), partials, &nb, stream);
    launch_reduce_final(partials, nb, dScalar + 1, stream, hScalarDev + 1);
}

// normFactors terms with y = gAverage(x) (solver.C:35-46), host slot 2.
void bmg_ctx::enqueueNormTerms(double y)
{
    Level& q0 = lv[0];
    int nb = 0;
    launch_norm_factor_terms<double>(q0.L, op<double>(0), cpr(q0.r),
                                     cpr(q0.b), y, partials, &nb, stream);
    launch_reduce_final(partials, nb, dScalar + 2, stream, hScalarDev + 2);
}

// x[0].correct<eliminatedBcs>() (MG.C:475)
void bmg_ctx::enqueueSolveEnd()
{
    launch_bc_fill(lv[0].L, op<double>(0), pr(lv[0].x), 2, 1, &bcval, stream);
}

// MG::Solve (MG.C:132-254) around the cycles; FP64 or mixed precision.
void bmg_ctx::runSolve(bool constMatrix, bmg_stats* st)
{
    if (opts.krylov == BMG_KRYLOV_FCG || opts.krylov == BMG_KRYLOV_FCG_LEARNED)
    {
        runFCG(constMatrix, st);
        return;
    }
    prepare(constMatrix);
    const bool mixed = opts.precision == BMG_MIXED;
    if (mixed) ensureMixed();

    const double cells = double(lv[0].cells());
    enqueueSolveStart();
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(stream));
    const double ss0 = hScalar[0];
    enqueueNormTerms(hScalar[1]/cells);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(stream));
    const double nf = std::max(std::sqrt(hScalar[2]), cells*1e-12);
    const double init = std::sqrt(ss0)/nf;
    if (!std::isfinite(init))
        throw Error(BMG_ERR_NUMERIC, "initial residual is not finite");
    double cur = init;
    auto check = [&](double c) -> int
    {
        return (init > 0.0 && c > opts.tolerance &&
                c/(init + 1e-20) > opts.rel_tol) ? 0 : 1;
    };
    int conv = check(cur);
    int iter = 0;
    while ((iter < opts.max_iter && conv == 0) || iter < opts.min_iter)
    {
        if (iter < opts.min_iter) conv = 0;
        const double ss = iterate(mixed);
        cur = std::sqrt(ss)/nf;
        if (!std::isfinite(cur))
            throw Error(BMG_ERR_NUMERIC, "residual norm is not finite");
        conv = check(cur);
        iter++;
    }
    enqueueSolveEnd();
    HIPCK(hipGetLastError());
    if (st)
    {
        st->iters = iter;
        st->initial_residual = init;
        st->final_residual = cur;
    }
}

namespace {

int run(const std::function<void()>& f)
{
    try
    {
        f();
        return BMG_OK;
    }
    catch (const Error& e)
    {
        g_error = e.what();
        return e.code;
    }
    catch (const std::exception& e)
    {
        g_error = e.what();
        return BMG_ERR_ARG;
    }
    catch (...)
    {
        g_error = "unknown error";
        return BMG_ERR_ARG;
    }
}

void buildAxis(HostAxis& a, const std::vector<double>& verts, bool periodic,
               int axis)
{
    const int N = a.N;
    a.v.assign(N + 3, 0.0);
    for (int c = 0; c <= N; c++) a.v[c + 1] = verts[c];
    // levelPoints.C:230-237 (mirrored) / :389-401 (periodic)
    if (!periodic)
    {
        a.v[0] = a.v[1] + (a.v[1] - a.v[2]);
        a.v[N + 2] = a.v[N + 1] + (a.v[N + 1] - a.v[N]);
    }
    else
    {
        a.v[0] = a.v[1] - (a.v[N + 1] - a.v[N]);
        a.v[N + 2] = a.v[N + 1] + (a.v[2] - a.v[1]);
    }
    for (auto& s : a.v) s = trimPrecision(s);
    a.cc.assign(N + 2, 0.0);
    a.w.assign(N + 2, 0.0);
    a.dl.assign(N + 2, 0.0);
    for (int p = 0; p < N + 2; p++)
    {
        a.cc[p] = hexaCentre(a.v[p], a.v[p + 1], axis);
        a.w[p] = a.v[p + 1] - a.v[p];
    }
    // fvMeshMetrics.C:876
    for (int p = 1; p < N + 2; p++)
        a.dl[p] = 1.0/std::fabs(a.cc[p] - a.cc[p - 1]);
}

void checkLevel(const bmg_ctx* c, int level)
{
    if (!c) throw Error(BMG_ERR_ARG, "null context");
    if (level < 0 || level >= c->nLevels())
        throw Error(BMG_ERR_ARG, "level out of range");
}

}  // namespace

extern "C" {

void bmg_default_opts(bmg_opts* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->rel_tol = 0.0;
    o->tolerance = 0.0;
    o->min_iter = 0;
    o->max_iter = 999;
    o->n_sweeps_pre = 0;
    o->n_sweeps_post = 2;
    o->cycle = BMG_FCYCLE;
    o->coarse_mode = BMG_COARSE_DIRECT;
    o->coarse_level = 0;
    o->smoother = BMG_RBGS;
    o->precision = BMG_FP64;
}

const char* bmg_last_error(void) { return g_error.c_str(); }

int bmg_device_count(int* count)
{
    return run([&] {
        if (!count) throw Error(BMG_ERR_ARG, "null pointer");
        HIPCK(hipGetDeviceCount(count));
    });
}

static void validateOpts(const bmg_opts* o)
{
    if (o->max_iter < 0 || o->min_iter < 0 || o->n_sweeps_pre < 0 ||
        o->n_sweeps_post < 0 || o->coarse_level < 0)
        throw Error(BMG_ERR_ARG, "negative iteration/sweep/level option");
    if (o->cycle < 0 || o->cycle > 2 || o->coarse_mode < 0 ||
        o->coarse_mode > 1 || o->smoother < 0 || o->smoother > 2 ||
        o->precision < 0 || o->precision > 1 || o->krylov < 0 ||
        o->krylov > 2)
        throw Error(BMG_ERR_ARG, "unknown enum value in options");
}

int bmg_create(bmg_ctx** out, const bmg_grid* g, const bmg_opts* opts)
{
    return run([&] {
        if (!out || !g) throw Error(BMG_ERR_ARG, "null pointer");
        *out = nullptr;
        for (int d = 0; d < 3; d++)
        {
            if (g->n[d] < 1 || g->n[d] > (1 << 14) || !g->vertices[d])
                throw Error(BMG_ERR_ARG, "invalid grid shape or vertices");
            for (int q = 0; q < g->n[d]; q++)
                if (!(g->vertices[d][q + 1] > g->vertices[d][q]))
                    throw Error(BMG_ERR_ARG,
                                "vertex coordinates must increase");
        }
        for (int f = 0; f < 6; f++)
            if (g->bc[f] < 0 || g->bc[f] > 2)
                throw Error(BMG_ERR_ARG, "unknown boundary kind");
        for (int d = 0; d < 3; d++)
            if ((g->bc[2*d] == BMG_PERIODIC) != (g->bc[2*d + 1] == BMG_PERIODIC))
                throw Error(BMG_ERR_ARG,
                            "periodic faces must come in opposite pairs");
        std::unique_ptr<bmg_ctx> c(new bmg_ctx);
        if (const char* e = std::getenv("BMG_NO_GRAPHS"))
            c->useGraphs = e[0] == '0';
        HIPCK(hipGetDevice(&c->device));
        HIPCK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCK(hipEventCreate(&c->ev0));
        HIPCK(hipEventCreate(&c->ev1));
        for (int f = 0; f < 6; f++) { c->bc[f] = g->bc[f]; c->bcval.v[f] = g->bc_value[f]; }
        bmg_opts o;
        bmg_default_opts(&o);
        if (opts) o = *opts;
        validateOpts(&o);
        c->opts = o;

        // mesh::generateLevels (mesh.C:19-53), level::NandR (level.C:11-53)
        int N[3] = {g->n[0], g->n[1], g->n[2]};
        int R[3] = {0, 0, 0};
        while (true)
        {
            Level q;
            for (int d = 0; d < 3; d++)
            {
                q.L.N[d] = N[d];
                q.L.P[d] = N[d] + 2;
                q.L.R[d] = R[d];
                q.ax[d].N = N[d];
                q.ax[d].R = R[d];
                std::vector<double> v(N[d] + 1);
                if (c->lv.empty())
                    for (int i = 0; i <= N[d]; i++) v[i] = g->vertices[d][i];
                else
                {
                    const HostAxis& pa = c->lv.back().ax[d];
                    for (int i = 0; i <= N[d]; i++) v[i] = pa.v[i*R[d] + 1];
                }
                buildAxis(q.ax[d], v, c->periodic(2*d), d);
            }
            for (int f = 0; f < 6; f++) q.L.bc[f] = c->bc[f];
            q.L.Hz = (q.L.P[2] + 1)/2;
            q.L.sI = q.L.P[1]*q.L.Hz;
            q.L.ncol = long(q.L.P[0])*q.L.P[1]*q.L.Hz;
            if (q.L.ncol*8 >= (1L << 31))
                throw Error(BMG_ERR_ARG, "level too large for 32-bit colour "
                                         "indexing");
            c->lv.push_back(q);
            int Nc[3];
            bool coarsenable = false;
            for (int d = 0; d < 3; d++)
            {
                // utils.H:1156-1161 rejects these under FULLDEBUG
                if (N[d] > 3 && N[d] % 2)
                    throw Error(BMG_ERR_ARG, "non-coarsenable grid: a level "
                                             "has an odd size above 3");
                Nc[d] = N[d] <= 3 ? N[d] : N[d]/2;
                if (Nc[d] != N[d]) coarsenable = true;
            }
            if (!coarsenable) break;
            for (int d = 0; d < 3; d++) { R[d] = N[d]/Nc[d]; N[d] = Nc[d]; }
        }
        if (c->opts.coarse_level > c->nLevels() - 2 && c->nLevels() > 1)
            throw Error(BMG_ERR_ARG, "coarse level is set too high");

        // device storage
        size_t maxBlocks = 1024;
        for (auto& q : c->lv)
        {
            for (int cc = 0; cc < 2; cc++)
            {
                q.x[cc] = devAlloc<double>(q.L.ncol);
                q.b[cc] = devAlloc<double>(q.L.ncol);
                q.r[cc] = devAlloc<double>(q.L.ncol);
                HIPCK(hipMemset(q.x[cc], 0, q.L.ncol*sizeof(double)));
                HIPCK(hipMemset(q.b[cc], 0, q.L.ncol*sizeof(double)));
                HIPCK(hipMemset(q.r[cc], 0, q.L.ncol*sizeof(double)));
            }
            for (int d = 0; d < 3; d++)
            {
                q.dw[d] = devUpload(q.ax[d].w);
                q.ddl[d] = devUpload(q.ax[d].dl);
                if (c->periodic(2*d))
                    for (int k = 0; k < 2; k++)
                        q.snap[d][k] = devAlloc<double>(q.planeCells(d));
            }
            const size_t nbx = size_t((q.L.P[2] + 2)/2 + 1 + 63)/64 + 1;
            maxBlocks = std::max(maxBlocks,
                                 nbx*size_t(q.L.P[1]/4 + 2)*size_t(q.L.P[0]));
            maxBlocks = std::max(maxBlocks, size_t(pair_tile_count(q.L)));
        }
        // transfer weights (linearRestrictionScheme.C:15-80,
        // linearProlongationScheme.C:12-110), per axis for rectilinear meshes
        for (int l = 0; l + 1 < c->nLevels(); l++)
        {
            Level& f = c->lv[l];
            Level& q = c->lv[l + 1];
            for (int d = 0; d < 3; d++)
            {
                const HostAxis& fa = f.ax[d];
                const HostAxis& ca = q.ax[d];
                std::vector<double> ru(q.L.P[d], 0.0);
                for (int I = 1; I <= q.L.N[d]; I++)
                {
                    const int f0 = (I - 1)*q.L.R[d] + 1;
                    const double Px = ca.cc[I] - fa.cc[f0];
                    const double Dx = fa.cc[f0 + 1] - fa.cc[f0];
                    ru[I] = round12((Px*Dx)/(Dx*Dx));
                }
                q.ru[d] = devUpload(ru);
                std::vector<double> pu(f.L.P[d], 0.0);
                for (int I = 1; I <= f.L.N[d]; I++)
                {
                    const int i = I - 1;
                    const int off = (i % 2) ? 1 : -1;
                    const int base = i/q.L.R[d];
                    const double x0 = ca.cc[base + 1];
                    const double x1 = ca.cc[base + off + 1];
                    const double Px = fa.cc[I] - x0;
                    const double Dx = x1 - x0;
                    pu[I] = round12((Px*Dx)/(Dx*Dx));
                }
                f.pu[d] = devUpload(pu);
            }
        }
        // saved level-0 ghost coefficients of a stored stencil, one per
        // boundary cell of each non-periodic face
        for (int f = 0; f < 6; f++)
        {
            if (c->periodic(f)) continue;
            const int d = f/2;
            c->gsave[f] = devAlloc<double>(size_t(c->lv[0].L.N[d == 0 ? 1 : 0])*
                                           c->lv[0].L.N[d == 2 ? 1 : 2]);
        }
        c->npartials = 3*maxBlocks;
        c->partials = devAlloc<double>(3*maxBlocks);
        c->dScalar = devAlloc<double>(4);
        HIPCK(hipHostMalloc(&c->hScalar, 4*sizeof(double),
                            hipHostMallocMapped));
        HIPCK(hipHostGetDevicePointer((void**)&c->hScalarDev, c->hScalar, 0));
        *out = c.release();
    });
}

int bmg_destroy(bmg_ctx* c)
{
    return run([&] {
        if (c)
        {
            (void)hipStreamSynchronize(c->stream);
            delete c;
        }
    });
}

int bmg_set_opts(bmg_ctx* c, const bmg_opts* o)
{
    return run([&] {
        if (!c || !o) throw Error(BMG_ERR_ARG, "null pointer");
        validateOpts(o);
        if (o->coarse_level > c->nLevels() - 2 && c->nLevels() > 1)
            throw Error(BMG_ERR_ARG, "coarse level is set too high");
        if (o->coarse_level != c->opts.coarse_level ||
            o->coarse_mode != c->opts.coarse_mode)
            c->prepared = false;
        c->opts = *o;
        c->gen++;
    });
}

// Per-cell level-0 boundary data of one Neumann/Dirichlet face: the patch
// block boundaryValues_[0] / boundaryGradients_[0] of the face's boundary
// condition (DirichletBoundaryConditionBase.H:30,
// NeumannBoundaryConditionBase.H:30), copied to the device on the
// context's stream.  Captured graphs carry the data pointers, so a change
// from uniform to per-cell data (or back) invalidates them.
static void setBoundaryValues(bmg_ctx* c, int face, const double* v,
                              bool device)
{
    if (!c) throw Error(BMG_ERR_ARG, "null pointer");
    if (face < 0 || face > 5) throw Error(BMG_ERR_ARG, "face out of range");
    if (c->periodic(face))
        throw Error(BMG_ERR_ARG, "a periodic face has no boundary values");
    if (!v)
    {
        if (c->bcdev[face])
        {
            HIPCK(hipStreamSynchronize(c->stream));
            devFree(c->bcdev[face]);
            c->gen++;
        }
        c->bcval.p[face] = nullptr;
        return;
    }
    const int fd = face/2;
    size_t n = 1;
    for (int d = 0; d < 3; d++)
        if (d != fd) n *= size_t(c->lv[0].L.N[d]);
    if (!c->bcdev[face])
    {
        HIPCK(hipStreamSynchronize(c->stream));
        c->bcdev[face] = devAlloc<double>(n);
        c->gen++;
    }
    HIPCK(hipMemcpyAsync(c->bcdev[face], v, n*sizeof(double),
                         device ? hipMemcpyDeviceToDevice
                                : hipMemcpyHostToDevice, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    c->bcval.p[face] = c->bcdev[face];
}

int bmg_renew_source_terms(bmg_ctx* c)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        c->foldPending = true;
    });
}

int bmg_set_boundary_values(bmg_ctx* c, int face, const double* values)
{
    return run([&] { setBoundaryValues(c, face, values, false); });
}

int bmg_set_boundary_values_device(bmg_ctx* c, int face,
                                   const double* d_values)
{
    return run([&] { setBoundaryValues(c, face, d_values, true); });
}

int bmg_num_levels(const bmg_ctx* c) { return c ? c->nLevels() : BMG_ERR_ARG; }

int bmg_level_shape(const bmg_ctx* c, int level, int n[3])
{
    return run([&] {
        checkLevel(c, level);
        for (int d = 0; d < 3; d++) n[d] = c->lv[level].L.N[d];
    });
}

static void setStencil(bmg_ctx* c, int level, const double* A7, bool device,
                       int symmetric)
{
    checkLevel(c, level);
    if (!A7) throw Error(BMG_ERR_ARG, "null stencil");
    Level& q = c->lv[level];
    const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2]*7;
    double* d = nullptr;
    if (!device)
    {
        d = devAlloc<double>(n);
        HIPCK(hipMemcpyAsync(d, A7, n*sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    }
    for (int cc = 0; cc < 2; cc++)
    {
        devFree(q.wf[cc]);
        devFree(q.wff[cc]);
        devFree(q.dg[cc]);
        devFree(q.dgf[cc]);
        if (!q.a7[cc]) q.a7[cc] = devAlloc<double>(q.L.ncol*7);
        HIPCK(hipMemsetAsync(q.a7[cc], 0, q.L.ncol*7*sizeof(double),
                             c->stream));
    }
    launch_stencil_ref_to_split(q.L, device ? A7 : d, pr(q.a7), c->stream);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(c->stream));
    if (d) devFree(d);
    q.opkind = OP_STENCIL;
    q.haveOp = true;
    c->gen++;
    if (level == 0)
    {
        // a stored stencil carries no coefficient field to deflate by
        c->opGen++;
        c->deflLabel.clear();
        c->deflM = 0;
    }
    c->symmetric = symmetric != 0;
    c->eliminated = false;
    c->foldPending = true;
    c->prepared = false;
    c->mixedReady = false;
}

int bmg_set_operator_stencil(bmg_ctx* c, int level, const double* A7,
                             int symmetric)
{
    return run([&] { setStencil(c, level, A7, false, symmetric); });
}

int bmg_set_operator_stencil_device(bmg_ctx* c, int level, const double* A7,
                                    int symmetric)
{
    return run([&] { setStencil(c, level, A7, true, symmetric); });
}

// linearGaussLaplacianScheme::imLaplacian (linearGaussLaplacianScheme.C:
// 22-102) on every level: face coefficients area*delta*lambda (lambda
// restricted area-weighted to coarse levels) or the lambda-free operator;
// with dt > 0 minus im::ddt(x)/dtFrac of the Euler scheme, whose matrix is
// the diagonal V/dt on every level (EulerDdtScheme.C:50-54) divided by
// dtFrac (defaultPoissonSolver.C:91-92).
static void setLaplacian(bmg_ctx* c, const double* lx, const double* ly,
                         const double* lz, double dt, double dtFrac)
{
    {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        const bool withLambda = lx != nullptr;
        if (withLambda && (!ly || !lz))
            throw Error(BMG_ERR_ARG, "need all three lambda arrays");
        const bool withDdt = dt != 0.0;
        if (withDdt && !(std::isfinite(dt) && dt > 0.0 &&
                         std::isfinite(dtFrac) && dtFrac != 0.0))
            throw Error(BMG_ERR_ARG, "time term needs a finite delta_t > 0 "
                                     "and a finite, nonzero dt_frac");
        // input channel of the learned preconditioner: (log10(mean of the
        // six face coefficients) + 1.5) / 1.5 per cell, reference cell order
        {
            const LevelDims& L0 = c->lv[0].L;
            const int N0 = L0.N[0], N1 = L0.N[1], N2 = L0.N[2];
            const long P1 = L0.P[1], P2 = L0.P[2];
            c->coefChannel.assign(size_t(N0)*N1*N2, 1.0f);
            double logSum = 0.0;
            if (withLambda)
                for (int i = 1; i <= N0; i++)
                    for (int j = 1; j <= N1; j++)
                        for (int k = 1; k <= N2; k++)
                        {
                            const long p = (i*P1 + j)*P2 + k;
                            const double m = (lx[p] + lx[p + P1*P2] + ly[p] +
                                              ly[p + P2] + lz[p] + lz[p + 1])/6.0;
                            const double v = m > 0.0 ? std::log10(m) : -6.0;
                            const float ch = float((v + 1.5)/1.5);
                            c->coefChannel[(size_t(i - 1)*N1 + (j - 1))*N2 +
                                           (k - 1)] = ch;
                            logSum += 1.5*double(ch) - 1.5;
                        }
            c->coefScale = std::pow(10.0, logSum/double(c->coefChannel.size()));
            c->coefUploaded = false;
        }
        // deflation sets of the new operator (flexible CG)
        c->opGen++;
        if (withLambda && c->opts.deflation)
            c->deflationLabels(lx, ly, lz);
        else
        {
            c->deflLabel.clear();
            c->deflM = 0;
        }
        std::vector<double*> lam(3*c->nLevels(), nullptr);
        if (withLambda)
        {
            const Level& q0 = c->lv[0];
            const size_t n0 = size_t(q0.L.P[0])*q0.L.P[1]*q0.L.P[2];
            const double* h[3] = {lx, ly, lz};
            for (int fd = 0; fd < 3; fd++)
            {
                lam[fd] = devAlloc<double>(n0);
                HIPCK(hipMemcpyAsync(lam[fd], h[fd], n0*sizeof(double),
                                     hipMemcpyHostToDevice, c->stream));
            }
            for (int l = 0; l + 1 < c->nLevels(); l++)
            {
                const Level& f = c->lv[l];
                const Level& q = c->lv[l + 1];
                const size_t nc = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2];
                for (int fd = 0; fd < 3; fd++)
                {
                    lam[3*(l + 1) + fd] = devAlloc<double>(nc);
                    HIPCK(hipMemsetAsync(lam[3*(l + 1) + fd], 0,
                                         nc*sizeof(double), c->stream));
                    launch_lambda_restrict(f.L, q.L, f.dw[0], f.dw[1], f.dw[2],
                                           lam[3*l + fd], lam[3*(l + 1) + fd],
                                           fd, c->stream);
                }
            }
        }
        for (int l = 0; l < c->nLevels(); l++)
        {
            Level& q = c->lv[l];
            for (int cc = 0; cc < 2; cc++)
            {
                devFree(q.a7[cc]); devFree(q.a7f[cc]);
                devFree(q.wf[cc]); devFree(q.wff[cc]);
                devFree(q.dg[cc]); devFree(q.dgf[cc]);
            }
            if (withDdt)
            {
                // ghost entries stay zero; only interior cells are read
                for (int cc = 0; cc < 2; cc++)
                {
                    q.dg[cc] = devAlloc<double>(q.L.ncol);
                    HIPCK(hipMemsetAsync(q.dg[cc], 0,
                                         q.L.ncol*sizeof(double), c->stream));
                }
                launch_ddt_diagonal(q.L, q.dw[0], q.dw[1], q.dw[2], dt, dtFrac,
                                    pr(q.dg), c->stream);
            }
            if (withLambda)
            {
                for (int cc = 0; cc < 2; cc++)
                {
                    q.wf[cc] = devAlloc<double>(q.L.ncol*3);
                    HIPCK(hipMemsetAsync(q.wf[cc], 0,
                                         q.L.ncol*3*sizeof(double), c->stream));
                }
                for (int fd = 0; fd < 3; fd++)
                    launch_face_coeffs(q.L, q.dw[0], q.dw[1], q.dw[2],
                                       q.ddl[fd], lam[3*l + fd], fd,
                                       pr(q.wf), c->stream);
                q.opkind = OP_FACES;
            }
            else
                q.opkind = OP_CONST;
            if (q.cells() <= MATERIALIZE_CELLS)
            {
                for (int cc = 0; cc < 2; cc++)
                {
                    q.a7[cc] = devAlloc<double>(q.L.ncol*7);
                    HIPCK(hipMemsetAsync(q.a7[cc], 0,
                                         q.L.ncol*7*sizeof(double), c->stream));
                }
                launch_materialize_stencil(q.L, c->op<double>(l), pr(q.a7),
                                           c->stream);
                q.opkind = OP_STENCIL;
                HIPCK(hipStreamSynchronize(c->stream));
                for (int cc = 0; cc < 2; cc++)
                {
                    devFree(q.wf[cc]);
                    devFree(q.dg[cc]);
                }
            }
            q.haveOp = true;
        }
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
        for (double* p : lam) if (p) (void)hipFree(p);
        c->gen++;
        c->symmetric = true;
        c->eliminated = false;
        c->foldPending = true;
        c->prepared = false;
        c->mixedReady = false;
    }
}

int bmg_set_operator_laplacian(bmg_ctx* c, const double* lx, const double* ly,
                               const double* lz)
{
    return run([&] { setLaplacian(c, lx, ly, lz, 0.0, 1.0); });
}

int bmg_set_operator_laplacian_ddt(bmg_ctx* c, const double* lx,
                                   const double* ly, const double* lz,
                                   double delta_t, double dt_frac)
{
    return run([&] {
        if (!(delta_t > 0.0))
            throw Error(BMG_ERR_ARG, "delta_t must be > 0");
        setLaplacian(c, lx, ly, lz, delta_t, dt_frac);
    });
}

int bmg_get_operator_stencil(bmg_ctx* c, int level, double* A7)
{
    return run([&] {
        checkLevel(c, level);
        if (!A7) throw Error(BMG_ERR_ARG, "null output");
        Level& q = c->lv[level];
        const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2]*7;
        Pair<double> tmp{{nullptr, nullptr}};
        bool own = false;
        if (q.opkind != OP_STENCIL)
        {
            for (int cc = 0; cc < 2; cc++)
            {
                tmp.c[cc] = devAlloc<double>(q.L.ncol*7);
                HIPCK(hipMemsetAsync(tmp.c[cc], 0, q.L.ncol*7*sizeof(double),
                                     c->stream));
            }
            launch_materialize_stencil(q.L, c->op<double>(level), tmp,
                                       c->stream);
            own = true;
            if (c->eliminated) launch_eliminate(q.L, tmp, c->stream);
        }
        else
            tmp = pr(q.a7);
        double* d = devAlloc<double>(n);
        HIPCK(hipMemsetAsync(d, 0, n*sizeof(double), c->stream));
        launch_stencil_split_to_ref(q.L, Pair<const double>{{tmp.c[0], tmp.c[1]}},
                                    d, c->stream);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(A7, d, n*sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        devFree(d);
        if (own) { devFree(tmp.c[0]); devFree(tmp.c[1]); }
    });
}

static void solveImpl(bmg_ctx* c, double* b, double* x, bool device,
                      int constMatrix, bmg_stats* st)
{
    if (!c || !b || !x) throw Error(BMG_ERR_ARG, "null pointer");
    Level& q = c->lv[0];
    const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2];
    double* db = b;
    double* dx = x;
    if (!device)
    {
        db = devAlloc<double>(n);
        dx = devAlloc<double>(n);
        HIPCK(hipMemcpyAsync(db, b, n*sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
        HIPCK(hipMemcpyAsync(dx, x, n*sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    }
    bmg_stats local{};
    std::exception_ptr err;
    try
    {
        HIPCK(hipEventRecord(c->ev0, c->stream));
        launch_ref_to_split<double>(q.L, db, pr(q.b), c->stream);
        launch_ref_to_split<double>(q.L, dx, pr(q.x), c->stream);
        c->runSolve(constMatrix != 0, &local);
        launch_split_to_ref<double>(q.L, cpr(q.x), dx, c->stream);
        launch_split_to_ref<double>(q.L, cpr(q.b), db, c->stream);
        HIPCK(hipEventRecord(c->ev1, c->stream));
        HIPCK(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        HIPCK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        local.solve_ms = ms;
        if (!device)
        {
            HIPCK(hipMemcpyAsync(x, dx, n*sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
            HIPCK(hipMemcpyAsync(b, db, n*sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
            HIPCK(hipStreamSynchronize(c->stream));
        }
    }
    catch (...)
    {
        err = std::current_exception();
    }
    if (!device)
    {
        (void)hipStreamSynchronize(c->stream);
        devFree(db);
        devFree(dx);
    }
    if (err)
    {
        if (st)
        {
            *st = local;
            st->status = BMG_ERR_ARG;
            try { std::rethrow_exception(err); }
            catch (const Error& e) { st->status = e.code; }
            catch (...) {}
        }
        std::rethrow_exception(err);
    }
    local.status = BMG_OK;
    if (st) *st = local;
}

int bmg_solve(bmg_ctx* c, double* b, double* x, int constMatrix, bmg_stats* st)
{
    return run([&] { solveImpl(c, b, x, false, constMatrix, st); });
}

int bmg_fft_create(bmg_fft** out, const bmg_grid* grid)
{
    return run([&] { fft_create(out, grid); });
}

int bmg_fft_destroy(bmg_fft* f)
{
    return run([&] { fft_destroy(f); });
}

int bmg_fft_solve(bmg_fft* f, const double* b, double* x)
{
    return run([&] { fft_solve(f, b, x); });
}

int bmg_fft_solve_device(bmg_fft* f, const double* db, double* dx)
{
    return run([&] { fft_solve_device(f, db, dx); });
}

static double fftTimeStep(const double* xold, double delta_t, double dt_frac)
{
    if (!xold) throw Error(BMG_ERR_ARG, "the time term needs x_old");
    const double dT = delta_t*dt_frac;
    if (!(std::isfinite(delta_t) && std::isfinite(dt_frac) && dT > 0.0))
        throw Error(BMG_ERR_ARG, "the time term needs delta_t*dt_frac > 0");
    return dT;
}

int bmg_fft_solve_ddt(bmg_fft* f, const double* b, double* x,
                      const double* x_old, double delta_t, double dt_frac)
{
    return run([&] {
        fft_solve(f, b, x, x_old, fftTimeStep(x_old, delta_t, dt_frac));
    });
}

int bmg_fft_solve_ddt_device(bmg_fft* f, const double* db, double* dx,
                             const double* dx_old, double delta_t,
                             double dt_frac)
{
    return run([&] {
        fft_solve_device(f, db, dx, dx_old,
                         fftTimeStep(dx_old, delta_t, dt_frac));
    });
}

int bmg_batch_create(bmg_batch** out, bmg_ctx* const* ctxs, int n)
{
    return run([&] { batch_create(out, ctxs, n); });
}

int bmg_batch_destroy(bmg_batch* b)
{
    return run([&] { batch_destroy(b); });
}

int bmg_solve_batch(bmg_batch* b, double* const* db, double* const* dx,
                    int constMatrix, bmg_stats* st)
{
    return run([&] { batch_solve(b, db, dx, constMatrix != 0, st); });
}

int bmg_batch_info(const bmg_batch* b, int* launches, long* workgroups)
{
    return run([&] { batch_info(b, launches, workgroups); });
}

int bmg_deflation_vectors(const bmg_ctx* c, int* count)
{
    return run([&] {
        if (!c || !count) throw Error(BMG_ERR_ARG, "null pointer");
        *count = c->opts.deflation ? c->deflM : 0;
    });
}

int bmg_batch_tuning(const bmg_batch* b, int* launches, int* reshaped)
{
    return run([&] { batch_tuning(b, launches, reshaped); });
}

int bmg_batch_set_tile_shape(bmg_batch* b, int ws)
{
    return run([&] { batch_set_tile_shape(b, ws); });
}

int bmg_batch_set_autotune(bmg_batch* b, int enable)
{
    return run([&] { batch_set_autotune(b, enable); });
}

int bmg_solve_device(bmg_ctx* c, double* db, double* dx, int constMatrix,
                     bmg_stats* st)
{
    return run([&] { solveImpl(c, db, dx, true, constMatrix, st); });
}

// MG::solve of a diagonal system (MG.C:404-428): level-0 stencil, ghost
// elimination, x = b/A_C, level-0 boundary update.
int bmg_solve_diagonal(bmg_ctx* c, const double* A7, double* b, double* x)
{
    return run([&] {
        if (!c || !A7 || !b || !x) throw Error(BMG_ERR_ARG, "null pointer");
        setStencil(c, 0, A7, false, 1);
        Level& q = c->lv[0];
        const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2];
        double* db = devAlloc<double>(n);
        double* dx = devAlloc<double>(n);
        HIPCK(hipMemcpyAsync(db, b, n*sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
        HIPCK(hipMemcpyAsync(dx, x, n*sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
        launch_ref_to_split<double>(q.L, db, pr(q.b), c->stream);
        launch_ref_to_split<double>(q.L, dx, pr(q.x), c->stream);
        if (c->inhomogeneous())
            launch_bc_source(q.L, c->op<double>(0), pr(q.b), &c->bcval, c->stream);
        launch_eliminate(q.L, pr(q.a7), c->stream);
        c->eliminated = true;
        launch_diag(q.L, cpr(q.a7), pr(q.x), cpr(q.b), c->stream);
        launch_bc_fill(q.L, c->op<double>(0), pr(q.x), 1, 1, &c->bcval, c->stream);
        launch_split_to_ref<double>(q.L, cpr(q.x), dx, c->stream);
        launch_split_to_ref<double>(q.L, cpr(q.b), db, c->stream);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(x, dx, n*sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
        HIPCK(hipMemcpyAsync(b, db, n*sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        devFree(db);
        devFree(dx);
    });
}

// PoissonSolver flux (defaultPoissonSolver.C:108-121) of a level-0 field.
static void faceFlux(bmg_ctx* c, const double* x, const double* const lam[3],
                     double* const flux[3])
{
    Level& q = c->lv[0];
    const double* w[3] = {q.dw[0], q.dw[1], q.dw[2]};
    for (int d = 0; d < 3; d++)
        launch_face_flux(q.L, x, lam[d], w, q.ddl[d], d, flux[d], c->stream);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(c->stream));
}

// splitPoissonSolver's correction source (splitPoissonSolver.C:124-137) on
// level 0: phi per face direction, then bHat.
static void splitSource(bmg_ctx* c, const double* x, const double* const lam[3],
                        double lam0, const double* b, double* const phi[3],
                        double* bhat)
{
    if (!(std::isfinite(lam0) && lam0 > 0.0))
        throw Error(BMG_ERR_ARG, "lambda0 must be finite and > 0");
    Level& q = c->lv[0];
    const double* w[3] = {q.dw[0], q.dw[1], q.dw[2]};
    const double* dl[3] = {q.ddl[0], q.ddl[1], q.ddl[2]};
    launch_split_source(q.L, x, lam, lam0, b, w, dl, phi, bhat, c->stream);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(c->stream));
}

int bmg_split_source_device(bmg_ctx* c, const double* dx, const double* dlx,
                            const double* dly, const double* dlz,
                            double lambda0, const double* db, double* dbhat,
                            double* dphix, double* dphiy, double* dphiz)
{
    return run([&] {
        if (!c || !dx || !dlx || !dly || !dlz || !dbhat || !dphix || !dphiy ||
            !dphiz)
            throw Error(BMG_ERR_ARG, "null pointer");
        const double* lam[3] = {dlx, dly, dlz};
        double* const phi[3] = {dphix, dphiy, dphiz};
        splitSource(c, dx, lam, lambda0, db, phi, dbhat);
    });
}

int bmg_split_source(bmg_ctx* c, const double* x, const double* lx,
                     const double* ly, const double* lz, double lambda0,
                     const double* b, double* bhat, double* phix, double* phiy,
                     double* phiz)
{
    return run([&] {
        if (!c || !x || !lx || !ly || !lz || !bhat || !phix || !phiy || !phiz)
            throw Error(BMG_ERR_ARG, "null pointer");
        Level& q = c->lv[0];
        const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2];
        std::vector<double*> bufs;
        auto up = [&](const double* h) -> double* {
            double* d = devAlloc<double>(n);
            bufs.push_back(d);
            if (h)
                HIPCK(hipMemcpy(d, h, n*sizeof(double), hipMemcpyHostToDevice));
            else
                HIPCK(hipMemset(d, 0, n*sizeof(double)));
            return d;
        };
        try
        {
            const double* dx = up(x);
            const double* lam[3] = {up(lx), up(ly), up(lz)};
            const double* db = b ? up(b) : nullptr;
            double* dbh = up(bhat);
            double* const phi[3] = {up(phix), up(phiy), up(phiz)};
            splitSource(c, dx, lam, lambda0, db, phi, dbh);
            HIPCK(hipMemcpy(bhat, dbh, n*sizeof(double),
                            hipMemcpyDeviceToHost));
            double* const out[3] = {phix, phiy, phiz};
            for (int d = 0; d < 3; d++)
                HIPCK(hipMemcpy(out[d], phi[d], n*sizeof(double),
                                hipMemcpyDeviceToHost));
        }
        catch (...)
        {
            for (auto& p : bufs) devFree(p);
            throw;
        }
        for (auto& p : bufs) devFree(p);
    });
}

int bmg_face_flux_device(bmg_ctx* c, const double* dx, const double* dlx,
                         const double* dly, const double* dlz, double* dfx,
                         double* dfy, double* dfz)
{
    return run([&] {
        if (!c || !dx || !dfx || !dfy || !dfz)
            throw Error(BMG_ERR_ARG, "null pointer");
        if ((dlx || dly || dlz) && !(dlx && dly && dlz))
            throw Error(BMG_ERR_ARG, "give all three lambda arrays or none");
        const double* lam[3] = {dlx, dly, dlz};
        double* const flux[3] = {dfx, dfy, dfz};
        faceFlux(c, dx, lam, flux);
    });
}

int bmg_face_flux(bmg_ctx* c, const double* x, const double* lx,
                  const double* ly, const double* lz, double* fx, double* fy,
                  double* fz)
{
    return run([&] {
        if (!c || !x || !fx || !fy || !fz)
            throw Error(BMG_ERR_ARG, "null pointer");
        if ((lx || ly || lz) && !(lx && ly && lz))
            throw Error(BMG_ERR_ARG, "give all three lambda arrays or none");
        Level& q = c->lv[0];
        const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2];
        std::vector<double*> bufs;
        auto up = [&](const double* h) -> double* {
            double* d = devAlloc<double>(n);
            bufs.push_back(d);
            if (h)
                HIPCK(hipMemcpy(d, h, n*sizeof(double), hipMemcpyHostToDevice));
            else
                HIPCK(hipMemset(d, 0, n*sizeof(double)));
            return d;
        };
        try
        {
            const double* dx = up(x);
            const double* lam[3] = {lx ? up(lx) : nullptr, ly ? up(ly) : nullptr,
                                    lz ? up(lz) : nullptr};
            double* const flux[3] = {up(nullptr), up(nullptr), up(nullptr)};
            faceFlux(c, dx, lam, flux);
            double* const out[3] = {fx, fy, fz};
            for (int d = 0; d < 3; d++)
                HIPCK(hipMemcpy(out[d], flux[d], n*sizeof(double),
                                hipMemcpyDeviceToHost));
        }
        catch (...)
        {
            for (auto& p : bufs) devFree(p);
            throw;
        }
        for (auto& p : bufs) devFree(p);
    });
}

// Staging memory of bmg_face_flux_noalloc: a fixed array in the code
// object, one copy per device, held until the process ends.  Host code never
// passes the array's name to a kernel or a copy (in host code it names a
// host-side stand-in): it uses the device address hipGetSymbolAddress
// returns for the device, taken once per device with that device current.
static __device__ __align__(256) double scratch[BMG_NOALLOC_SCRATCH_DOUBLES];

static std::mutex g_scratchMutex;
static double* g_scratchPtr[64] = {};

static double* scratchOn(int device)
{
    if (device < 0 || device >= 64)
        throw Error(BMG_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[device])
    {
        int prev = 0;
        HIPCK(hipGetDevice(&prev));
        if (prev != device) HIPCK(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) HIPCK(hipSetDevice(prev));
        if (e != hipSuccess || !p)
            throw Error(BMG_ERR_HIP, std::string("hipGetSymbolAddress of the "
                                                 "staging array: ") +
                                         hipGetErrorString(e));
        g_scratchPtr[device] = static_cast<double*>(p);
    }
    return g_scratchPtr[device];
}

int bmg_face_flux_noalloc(bmg_ctx* c, const double* x, const double* lx,
                          const double* ly, const double* lz, double* fx,
                          double* fy, double* fz)
{
    return run([&] {
        if (!c || !x || !fx || !fy || !fz)
            throw Error(BMG_ERR_ARG, "null pointer");
        const bool lam = lx || ly || lz;
        if (lam && !(lx && ly && lz))
            throw Error(BMG_ERR_ARG, "give all three lambda arrays or none");
        Level& q = c->lv[0];
        const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2];
        const size_t need = n*(lam ? 7 : 4);
        if (need > size_t(BMG_NOALLOC_SCRATCH_DOUBLES))
            throw Error(BMG_ERR_ARG,
                        "level 0 of " + std::to_string(n) + " padded cells "
                        "needs " + std::to_string(need) + " doubles of "
                        "staging, more than the " +
                        std::to_string(BMG_NOALLOC_SCRATCH_DOUBLES) +
                        " of bmg_face_flux_noalloc");
        double* s = scratchOn(c->device);
        // staging: x, the three fluxes, then the three lambda arrays
        double* dx = s;
        double* const flux[3] = {s + n, s + 2*n, s + 3*n};
        const double* dl[3] = {nullptr, nullptr, nullptr};
        HIPCK(hipMemcpyAsync(dx, x, n*sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
        // entries outside the face ranges are returned as 0; the array keeps
        // the previous call's values, so clear the fluxes every call
        HIPCK(hipMemsetAsync(flux[0], 0, 3*n*sizeof(double), c->stream));
        if (lam)
        {
            const double* h[3] = {lx, ly, lz};
            for (int d = 0; d < 3; d++)
            {
                double* p = s + (4 + d)*n;
                HIPCK(hipMemcpyAsync(p, h[d], n*sizeof(double),
                                     hipMemcpyHostToDevice, c->stream));
                dl[d] = p;
            }
        }
        faceFlux(c, dx, dl, flux);
        double* const out[3] = {fx, fy, fz};
        for (int d = 0; d < 3; d++)
            HIPCK(hipMemcpyAsync(out[d], flux[d], n*sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

static double* fieldPtr(bmg_ctx* c, int which, int level, int colour)
{
    Level& q = c->lv[level];
    switch (which)
    {
    case 0: return q.x[colour];
    case 1: return q.b[colour];
    case 2: return q.r[colour];
    default: throw Error(BMG_ERR_ARG, "unknown field");
    }
}

int bmg_upload_field(bmg_ctx* c, int which, int level, const double* v)
{
    return run([&] {
        checkLevel(c, level);
        if (!v) throw Error(BMG_ERR_ARG, "null pointer");
        Level& q = c->lv[level];
        const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2];
        double* d = devAlloc<double>(n);
        HIPCK(hipMemcpyAsync(d, v, n*sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
        Pair<double> f{{fieldPtr(c, which, level, 0),
                        fieldPtr(c, which, level, 1)}};
        launch_ref_to_split<double>(q.L, d, f, c->stream);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
        devFree(d);
    });
}

int bmg_download_field(bmg_ctx* c, int which, int level, double* v)
{
    return run([&] {
        checkLevel(c, level);
        if (!v) throw Error(BMG_ERR_ARG, "null pointer");
        Level& q = c->lv[level];
        const size_t n = size_t(q.L.P[0])*q.L.P[1]*q.L.P[2];
        double* d = devAlloc<double>(n);
        Pair<const double> f{{fieldPtr(c, which, level, 0),
                              fieldPtr(c, which, level, 1)}};
        launch_split_to_ref<double>(q.L, f, d, c->stream);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(v, d, n*sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        devFree(d);
    });
}

int bmg_op_prepare(bmg_ctx* c)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        c->prepare(false);
        c->enqueueBoundaryTerms();
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

static void needEliminated(bmg_ctx* c)
{
    if (!c->eliminated)
        throw Error(BMG_ERR_STATE, "call bmg_op_prepare first");
}

int bmg_op_smooth(bmg_ctx* c, int level, int sweeps)
{
    return run([&] {
        checkLevel(c, level);
        needEliminated(c);
        c->smooth<double>(level, sweeps);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int bmg_op_residual(bmg_ctx* c, int level, int corr)
{
    return run([&] {
        checkLevel(c, level);
        needEliminated(c);
        if (corr)
            c->residual<double>(level, true, false, nullptr);
        else
        {
            Level& q = c->lv[level];
            launch_residual(q.L, c->op<double>(level), cpr(q.x), cpr(q.b),
                            pr(q.r), nullptr, nullptr, c->stream);
        }
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int bmg_op_restrict(bmg_ctx* c, int level)
{
    return run([&] {
        checkLevel(c, level);
        if (level + 1 >= c->nLevels()) throw Error(BMG_ERR_ARG, "no coarser level");
        c->restrictTo<double>(level);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int bmg_op_prolong(bmg_ctx* c, int level)
{
    return run([&] {
        checkLevel(c, level);
        if (level + 1 >= c->nLevels()) throw Error(BMG_ERR_ARG, "no coarser level");
        if (!c->lv[level + 1].haveOp)
            throw Error(BMG_ERR_STATE, "operator not set");
        c->prolongFrom<double>(level);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int bmg_op_correct_bcs(bmg_ctx* c, int level)
{
    return run([&] {
        checkLevel(c, level);
        launch_bc_fill(c->lv[level].L, c->op<double>(level),
                       pr(c->lv[level].x), 1, level == 0, &c->bcval, c->stream);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int bmg_op_coarse_solve(bmg_ctx* c)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        needEliminated(c);
        c->coarseSolve<double>();
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int bmg_op_cycle(bmg_ctx* c)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        needEliminated(c);
        std::vector<int> visits(c->nLevels(), 0);
        c->cycle<double>(visits, 0);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int bmg_op_sum_sq_residual(bmg_ctx* c, double* out)
{
    return run([&] {
        if (!c || !out) throw Error(BMG_ERR_ARG, "null pointer");
        int nb = 0;
        launch_sum_sq<double>(c->lv[0].L, cpr(c->lv[0].r), c->partials, &nb,
                              c->stream);
        *out = c->reduceFinal(nb);
    });
}

static double phaseBytesPerUpdate(const Level& q, size_t w);

int bmg_time_smoother(bmg_ctx* c, int level, int precision, int reps,
                      double* avg_ms, double* bytes)
{
    return run([&] {
        checkLevel(c, level);
        needEliminated(c);
        if (reps < 1) throw Error(BMG_ERR_ARG, "reps must be positive");
        if (precision == BMG_MIXED) c->ensureMixed();
        Level& q = c->lv[level];
        const Region rg = c->interior(level, 1);
        HIPCK(hipStreamSynchronize(c->stream));
        HIPCK(hipEventRecord(c->ev0, c->stream));
        for (int r = 0; r < reps; r++)
        {
            if (precision == BMG_MIXED)
                launch_rbgs<float>(q.L, c->op<float>(level), pr(q.xf), cpr(q.bf),
                                   r & 1, rg, c->stream);
            else
                launch_rbgs<double>(q.L, c->op<double>(level), pr(q.x), cpr(q.b),
                                    r & 1, rg, c->stream);
        }
        HIPCK(hipEventRecord(c->ev1, c->stream));
        HIPCK(hipEventSynchronize(c->ev1));
        HIPCK(hipGetLastError());
        float ms = 0.0f;
        HIPCK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        long m[3];
        for (int d = 0; d < 3; d++) m[d] = std::max(0, rg.E[d] - rg.S[d]);
        const double cells = 0.5*double(m[0]*m[1]*m[2]);
        const int kind = q.opkind;
        (void)kind;
        if (avg_ms) *avg_ms = double(ms)/reps;
        if (bytes)
            *bytes = cells*phaseBytesPerUpdate(q, precision == BMG_MIXED
                                                      ? sizeof(float)
                                                      : sizeof(double));
    });
}

// Algorithmic bytes of the launches of one MG iteration (enqueueIteration):
// every array a launch must touch counted once (n interior cells of a level,
// w bytes per value):
//   red-black phase      n/2 (3w + coefficients of the representation)
//   residual             n (3w + coefficients), the norm fused
//   restriction          n_f w + n_c w;  prolongation  2 n_f w + n_c w
//   zeroing x            n w;  ghost fill  2 w per ghost cell written
//   periodic slabs       2 planes x (3w + coefficients) per face pair
//   coarse solve         the LU factors (8 n_c^2) + b in and x out
// plus, for mixed precision, the FP64 -> FP32 residual copy (12 n), the
// zeroing of e (4 n) and x += e (20 n) on level 0.
// Face coefficients are stored in FP16 for the FP32 cycles.
static double faceBytes(size_t w) { return w == sizeof(float) ? 2.0 : double(w); }

static double phaseBytesPerUpdate(const Level& q, size_t w)
{
    double per = 3.0*double(w);
    if (q.opkind == OP_FACES) per += 6.0*faceBytes(w);
    else if (q.opkind == OP_STENCIL) per += 7.0*double(w);
    if (q.opkind != OP_STENCIL && q.dg[0]) per += double(w);
    return per;
}

static double cellBytesFull(const Level& q, size_t w)
{
    // residual-type pass over every cell: x, b, r + each stored coefficient
    double per = 3.0*double(w);
    if (q.opkind == OP_FACES) per += 3.0*faceBytes(w);
    else if (q.opkind == OP_STENCIL) per += 7.0*double(w);
    if (q.opkind != OP_STENCIL && q.dg[0]) per += double(w);
    return per;
}

static double ghostCells(const Level& q)
{
    const double p0 = q.L.P[0], p1 = q.L.P[1], p2 = q.L.P[2];
    return p0*p1*p2 - double(q.cells());
}

static double smoothBytes(const bmg_ctx* c, int l, int sweeps, size_t w)
{
    const Level& q = c->lv[l];
    const Region rg = c->interior(l, 1);
    double cells = 1.0;
    for (int d = 0; d < 3; d++) cells *= std::max(0, rg.E[d] - rg.S[d]);
    double bytes = 0.0;
    int prev = -1;
    for (int s = 0; s < sweeps; s++)
    {
        int cols[4];
        const int nc = phase_colours(rg, cols);
        for (int i = 0; i < nc; i++)
        {
            if (cols[i] != prev) bytes += 0.5*cells*phaseBytesPerUpdate(q, w);
            prev = cols[i];
        }
        if (!c->anyPeriodic()) continue;
        bytes += 2.0*ghostCells(q)*double(w);
        for (int d = 0; d < 3; d++)
            if (c->periodic(2*d))
                bytes += 2.0*double(q.planeCells(d))*
                         (phaseBytesPerUpdate(q, w) + double(w));
    }
    return bytes;
}

static double cycleBytes(const bmg_ctx* c, int l, size_t w, bool zeroX)
{
    const Level& q = c->lv[l];
    const int lc = c->coarseLevel();
    const double n = double(q.cells());
    const bool direct = c->opts.coarse_mode == BMG_COARSE_DIRECT;
    const bool pre = c->opts.n_sweeps_pre > 0 && !(l == lc && direct);
    double bytes = 0.0;
    if (l > 0 || zeroX)
    {
        // x = 0: the FP32 cycles skip the fill where the first sweep
        // rewrites every cell, and that first phase reads no neighbour;
        // elsewhere a fill (or the zeroing first phase, same bytes less the
        // neighbour reads)
        const bool fp32 = w == sizeof(float);
        const bool skip = pre && fp32 && c->firstPhaseRewrites<float>(l);
        const bool fused = fp32 ? c->fusedLevel<float>(l)
                                : c->fusedLevel<double>(l);
        const bool zp = pre && !skip && c->opts.smoother == BMG_RBGS &&
                        !fused && n <= double(1L << 18);
        if (!skip) bytes += n*double(w);
        if (skip || zp) bytes -= 0.5*n*double(w);
    }
    if (pre)
        bytes += smoothBytes(c, l, c->opts.n_sweeps_pre, w);
    if (l < lc)
    {
        const Level& qc = c->lv[l + 1];
        const double nc = double(qc.cells());
        if (l > 0) bytes += n*cellBytesFull(q, w);
        bytes += n*double(w) + nc*double(w);
        bytes += cycleBytes(c, l + 1, w, true);
        bytes += 2.0*ghostCells(qc)*double(w);
        // prolongation: the fine cells read and written, half of them when
        // the first post-smoothing phase rewrites the other colour
        const bool by2 = q.L.N[0] == 2*qc.L.N[0] && q.L.N[1] == 2*qc.L.N[1] &&
                         q.L.N[2] == 2*qc.L.N[2];
        const int skip = w == sizeof(float) ? c->postPhaseColour<float>(l)
                                            : c->postPhaseColour<double>(l);
        bytes += (by2 && skip >= 0 ? 1.0 : 2.0)*n*double(w) + nc*double(w);
        if (c->opts.n_sweeps_post > 0)
            bytes += smoothBytes(c, l, c->opts.n_sweeps_post, w);
    }
    else if (direct)
        bytes += 8.0*n*n + 2.0*n*double(w) + 2.0*ghostCells(q)*double(w);
    else
        bytes += smoothBytes(c, l, std::max(c->opts.n_sweeps_post, 2), w);
    return bytes;
}

static double iterationBytes(const bmg_ctx* c, bool mixed)
{
    const Level& q0 = c->lv[0];
    const double n = double(q0.cells());
    // the FP32 cycle solves for a correction from 0 (its zeroing counted in
    // cycleBytes); the FP64 cycle starts from the iterate
    double bytes = cycleBytes(c, 0, mixed ? sizeof(float) : sizeof(double),
                              mixed);
    // r to FP32 (8 B read, 4 written), x += e (x read and written, e read)
    if (mixed) bytes += (12.0 + 20.0)*n;
    bytes += n*cellBytesFull(q0, sizeof(double));
    return bytes;
}

int bmg_time_op(bmg_ctx* c, int op, int level, int precision, int reps,
                double* avg_ms, double* bytes)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        checkLevel(c, level);
        needEliminated(c);
        if (reps < 1) throw Error(BMG_ERR_ARG, "reps must be positive");
        if (precision != BMG_FP64 && precision != BMG_MIXED)
            throw Error(BMG_ERR_ARG, "precision must be BMG_FP64 or BMG_MIXED");
        const bool mixed = precision == BMG_MIXED;
        if (mixed) c->ensureMixed();
        const size_t w = mixed ? sizeof(float) : sizeof(double);
        Level& q = c->lv[level];
        const double n = double(q.cells());
        const bool coarser = level + 1 < c->nLevels();
        if ((op == BMG_TIME_RESTRICT || op == BMG_TIME_PROLONG) && !coarser)
            throw Error(BMG_ERR_ARG, "no coarser level");
        if (op == BMG_TIME_ITERATION && level != 0)
            throw Error(BMG_ERR_ARG, "an iteration starts on level 0");
        if (op == BMG_TIME_ITERATION && !c->prepared)
            throw Error(BMG_ERR_STATE, "call bmg_op_prepare first");
        double by = 0.0;
        auto launchOnce = [&](int r) {
            switch (op)
            {
            case BMG_TIME_PHASE:
                if (mixed)
                    launch_rbgs<float>(q.L, c->op<float>(level), pr(q.xf),
                                       cpr(q.bf), r & 1, c->interior(level, 1),
                                       c->stream);
                else
                    launch_rbgs<double>(q.L, c->op<double>(level), pr(q.x),
                                        cpr(q.b), r & 1, c->interior(level, 1),
                                        c->stream);
                break;
            case BMG_TIME_RESIDUAL:
            {
                int nb = 0;
                if (mixed)
                    launch_residual(q.L, c->op<float>(level), cpr(q.xf),
                                    cpr(q.bf), pr(q.rf), nullptr, nullptr,
                                    c->stream);
                else
                    launch_residual(q.L, c->op<double>(level), cpr(q.x),
                                    cpr(q.b), pr(q.r), c->partials, &nb,
                                    c->stream);
                break;
            }
            case BMG_TIME_RESTRICT:
                if (mixed) c->restrictTo<float>(level);
                else c->restrictTo<double>(level);
                break;
            case BMG_TIME_PROLONG:
            {
                Level& qc = c->lv[level + 1];
                if (mixed)
                    launch_prolong(q.L, qc.L, c->transfer<float>(level),
                                   cpr(qc.xf), pr(q.xf), c->stream);
                else
                    launch_prolong(q.L, qc.L, c->transfer<double>(level),
                                   cpr(qc.x), pr(q.x), c->stream);
                break;
            }
            case BMG_TIME_ITERATION:
                HIPCK(hipGraphLaunch(c->graphExec, c->stream));
                break;
            default:
                throw Error(BMG_ERR_ARG, "unknown operation to time");
            }
        };
        switch (op)
        {
        case BMG_TIME_PHASE:
        {
            const Region rg = c->interior(level, 1);
            double cells = 1.0;
            for (int d = 0; d < 3; d++) cells *= std::max(0, rg.E[d] - rg.S[d]);
            by = 0.5*cells*phaseBytesPerUpdate(q, w);
            break;
        }
        case BMG_TIME_RESIDUAL:
            by = n*cellBytesFull(q, w);
            break;
        case BMG_TIME_RESTRICT:
            by = n*double(w) + double(c->lv[level + 1].cells())*double(w);
            break;
        case BMG_TIME_PROLONG:
            by = 2.0*n*double(w) + double(c->lv[level + 1].cells())*double(w);
            break;
        case BMG_TIME_ITERATION:
        {
            const bool keep = c->useGraphs;
            c->useGraphs = true;
            try { (void)c->iterate(mixed); }
            catch (...) { c->useGraphs = keep; throw; }
            c->useGraphs = keep;
            by = iterationBytes(c, mixed);
            break;
        }
        default:
            throw Error(BMG_ERR_ARG, "unknown operation to time");
        }
        launchOnce(0);   // warm caches and code
        HIPCK(hipStreamSynchronize(c->stream));
        HIPCK(hipEventRecord(c->ev0, c->stream));
        for (int r = 0; r < reps; r++) launchOnce(r);
        HIPCK(hipEventRecord(c->ev1, c->stream));
        HIPCK(hipEventSynchronize(c->ev1));
        HIPCK(hipGetLastError());
        float ms = 0.0f;
        HIPCK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (avg_ms) *avg_ms = double(ms)/reps;
        if (bytes) *bytes = by;
    });
}

int bmg_set_learned(bmg_ctx* c, bmg_unet* net)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        c->unet = net;
        c->gen++;   // the captured learned iteration refers to the network
    });
}

int bmg_set_graphs(bmg_ctx* c, int enable)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        c->useGraphs = enable != 0;
    });
}

int bmg_set_periodic_smoothing(bmg_ctx* c, int mode)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        if (mode < 0 || mode > 2)
            throw Error(BMG_ERR_ARG,
                        "periodic smoothing mode must be 0, 1 or 2");
        if (c->wrapSmoothing != mode) c->gen++;
        c->wrapSmoothing = mode;
    });
}

int bmg_set_fused_smoothing(bmg_ctx* c, int mode)
{
    return run([&] {
        if (!c) throw Error(BMG_ERR_ARG, "null context");
        if (mode < 0 || mode > 2)
            throw Error(BMG_ERR_ARG, "fused smoothing mode must be 0, 1 or 2");
        if (c->fuseSmoothing != mode) c->gen++;
        c->fuseSmoothing = mode;
    });
}

}  // extern "C"
