/*
 * bmg_oracle.cpp -- CPU restatement of Briscola's MG solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP solver and the "port" CPU baseline timed next to it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Built with -ffp-contract=off so every product and sum is rounded exactly as
 * written, in the order the reference writes it.  All citations are relative
 * to the reference repository root.
 *
 * Restatement choices (the reference itself cannot be built here: it needs
 * OpenFOAM-12):
 *   - meshes are rectilinear, given by per-axis vertex coordinates; coarse
 *     vertices are the fine ones sub-sampled (levelPoints.C:150-160), ghost
 *     vertices are mirrored (levelPoints.C:230-237) or, on periodic faces,
 *     continue with the opposite cell's width (levelPoints.C:389-401), and
 *     every vertex is trimmed to 14 significant digits (levelPoints.C:519-525,
 *     utils.H:1010-1021);
 *   - cell centres are the 8-vertex average in hexa::center's order
 *     (hexa.H:120-128); volumes are (dx*dy)*dz and face areas the product of
 *     the two in-face widths (the reference tessellates the hexahedron; for a
 *     box the two agree to rounding);
 *   - the transfer weights use the exact rectilinear initial guess of
 *     interpolationWeights (vertexVector.C:130-135) rounded to 1e-12
 *     (vertexVector.C:183-188); the one Newton correction the reference then
 *     applies moves u by O(1 ulp) before that rounding;
 *   - boundary conditions are uniform per face; OpenFOAM's LUDecompose is
 *     restated as Crout LU with implicit partial pivoting (Press et al.,
 *     Numerical Recipes, ludcmp/lubksb), which is the algorithm OpenFOAM-12
 *     src/OpenFOAM/matrices/scalarMatrices implements.
 */
#include "bmg_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

constexpr int NEUMANN = ORACLE_NEUMANN;
constexpr int DIRICHLET = ORACLE_DIRICHLET;
constexpr int PERIODIC = ORACLE_PERIODIC;

/* utils.H:1010-1021 */
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

/* vertexVector.C:183-188 */
double round12(double u)
{
    return std::round(u/1e-12)*1e-12;
}

struct Axis
{
    int N = 0;                 // interior cells
    int R = 0;                 // refinement ratio vs parent (0 on level 0)
    std::vector<double> v;     // vertices -1..N+1 stored at v[c+1]
    std::vector<double> cc;    // centre of padded cell p (= cell p-1)
    std::vector<double> w;     // width of padded cell p
    std::vector<double> dl;    // 1/|cc[p]-cc[p-1]| for the lower face of p
};

struct Level
{
    int N[3], P[3], R[3];
    Axis ax[3];
    size_t size() const { return size_t(P[0])*P[1]*P[2]; }
    long lin(int i, int j, int k) const
    {
        return (long(i)*P[1] + j)*P[2] + k;
    }
    std::vector<double> x, b, r, A;   // A: 7 per cell
    std::vector<double> lam[3];       // lower-face lambda (var-coef only)
    std::vector<double> rw;           // restriction weights (8 per cell)
    // prolongation weights and coarse indices (8 per interior fine cell)
    std::vector<double> pw;
    std::vector<long> pidx;
};

/* stencil.C:13-23 offsets for components 1..6 */
const int OFF[7][3] =
{
    {0,0,0}, {-1,0,0}, {1,0,0}, {0,-1,0}, {0,1,0}, {0,0,-1}, {0,0,1}
};

} // namespace

struct oracle_ctx
{
    std::vector<Level> L;
    int bc[6];
    double bcval[6];
    // per-cell level-0 boundary data of face f (Dirichlet value / Neumann
    // gradient), the reference's patch block boundaryValues_[d] /
    // boundaryGradients_[d] (DirichletBoundaryConditionBase.H:30,
    // NeumannBoundaryConditionBase.H:30): shape E - S of the face's layer,
    // index (ijk - S), last axis fastest; empty = uniform bcval[f]
    std::vector<double> bcarr[6];

    // operation counts of the last cycle: per level, MG::cycle calls
    // (visits), Smooth calls, residuals, restrictions, prolongations and
    // coarse solves
    std::vector<long> cnt[6];

    double relTol = 0, tolerance = 0;
    int minIter = 0, maxIter = 999, nPre = 0, nPost = 2;
    int cycle = ORACLE_FCYCLE, coarseMode = ORACLE_COARSE_DIRECT;
    int coarseLevel_ = 0, smoother = ORACLE_RBGS;

    bool symmetric = true;
    bool eliminated = false;

    // coarse direct solver (APLU)
    bool prepared = false;
    int nc = 0;
    std::vector<double> D, LU;
    std::vector<int> piv;

    int nLevels() const { return int(L.size()); }
    int coarseLevel() const { return nLevels() - coarseLevel_ - 1; }

    bool periodicFace(int f) const { return bc[f] == PERIODIC; }
    /* meshLevel.C:555-582: singular iff no Dirichlet face */
    bool singular() const
    {
        if (!symmetric) return false;
        for (int f = 0; f < 6; f++)
            if (bc[f] == DIRICHLET) return false;
        return true;
    }
};

namespace {

void buildAxis(Axis& a, const std::vector<double>& verts, bool periodic)
{
    const int N = a.N;
    a.v.assign(N + 3, 0.0);
    for (int c = 0; c <= N; c++) a.v[c+1] = verts[c];
    if (!periodic)
    {
        a.v[0] = a.v[1] + (a.v[1] - a.v[2]);
        a.v[N+2] = a.v[N+1] + (a.v[N+1] - a.v[N]);
    }
    else
    {
        a.v[0] = a.v[1] - (a.v[N+1] - a.v[N]);
        a.v[N+2] = a.v[N+1] + (a.v[2] - a.v[1]);
    }
    for (auto& s : a.v) s = trimPrecision(s);
}

/* hexa.H:120-128 centre, one axis: vertex pattern along this axis in
 * lba,rba,lta,rta,lbf,rbf,ltf,rtf order. */
double hexaCentre(double lo, double hi, int axis)
{
    static const int pat[3][8] =
    {
        {0,1,0,1,0,1,0,1},
        {0,0,1,1,0,0,1,1},
        {0,0,0,0,1,1,1,1}
    };
    double c = (pat[axis][0] ? hi : lo)/8.0;
    for (int i = 1; i < 8; i++)
        c += (pat[axis][i] ? hi : lo)/8.0;
    return c;
}

void axisMetrics(Axis& a, int axis)
{
    const int N = a.N;
    a.cc.assign(N + 2, 0.0);
    a.w.assign(N + 2, 0.0);
    a.dl.assign(N + 2, 0.0);
    for (int p = 0; p < N + 2; p++)
    {
        a.cc[p] = hexaCentre(a.v[p], a.v[p+1], axis);
        a.w[p] = a.v[p+1] - a.v[p];
    }
    /* fvMeshMetrics.C:846-893: 1/|cc_P - cc_N|; the other two centre
     * components are equal for rectilinear cells, so |.| = |dx|. */
    for (int p = 1; p < N + 2; p++)
        a.dl[p] = 1.0/std::fabs(a.cc[p] - a.cc[p-1]);
}

inline double& at(std::vector<double>& f, const Level& lv, int i, int j, int k)
{
    return f[lv.lin(i, j, k)];
}

/* Boundary-layer cell range [S,E) (padded indices) for offset bo. */
void layer(const Level& lv, const int bo[3], int S[3], int E[3])
{
    for (int d = 0; d < 3; d++)
    {
        if (bo[d] < 0) { S[d] = 1; E[d] = 2; }
        else if (bo[d] > 0) { S[d] = lv.N[d]; E[d] = lv.N[d] + 1; }
        else { S[d] = 1; E[d] = lv.N[d] + 1; }
    }
}

/* periodicBoundaryCondition.C:54-92 for one boundary offset bo: ghost cells
 * beyond the layer take the value of the cell one period away. */
void periodicCopy(const oracle_ctx& c, const Level& lv, std::vector<double>& f,
                  const int bo[3])
{
    int S[3], E[3];
    layer(lv, bo, S, E);
    for (int i = S[0]; i < E[0]; i++)
    for (int j = S[1]; j < E[1]; j++)
    for (int k = S[2]; k < E[2]; k++)
    {
        const int g[3] = {i + bo[0], j + bo[1], k + bo[2]};
        int s[3];
        for (int d = 0; d < 3; d++)
            s[d] = g[d] - bo[d]*lv.N[d];
        f[lv.lin(g[0], g[1], g[2])] = f[lv.lin(s[0], s[1], s[2])];
    }
    (void)c;
}

bool isFaceOffset(const int bo[3])
{
    return std::abs(bo[0]) + std::abs(bo[1]) + std::abs(bo[2]) == 1;
}

/* Does boundary offset bo carry a boundary condition?  Every face does; an
 * edge or corner does only when each face it borders is periodic (a
 * neighbour exists across it). */
bool hasBoundary(const oracle_ctx& c, const int bo[3])
{
    if (isFaceOffset(bo)) return true;
    for (int d = 0; d < 3; d++)
        if (bo[d] != 0 && !c.periodicFace(2*d + (bo[d] > 0)))
            return false;
    return true;
}

/* Boundary value of face f at boundary cell (i,j,k) of the layer [S,E):
 * boundaryValues_[d](ijk - S) (DirichletBoundaryCondition.C:45,68,
 * NeumannBoundaryCondition.C:51,83), or the uniform value. */
double bcValueAt(const oracle_ctx& c, int face, const int S[3], const int E[3],
                 int i, int j, int k)
{
    const std::vector<double>& a = c.bcarr[face];
    if (a.empty()) return c.bcval[face];
    const long n1 = E[1] - S[1], n2 = E[2] - S[2];
    return a[((long)(i - S[0])*n1 + (j - S[1]))*n2 + (k - S[2])];
}

/* Neumann/Dirichlet evaluate (NeumannBoundaryCondition.C:58-85,
 * DirichletBoundaryCondition.C:50-71) for face f of level l. */
void patchEvaluate(const oracle_ctx& c, int l, Level& lv,
                   std::vector<double>& f, int face)
{
    const int bo[3] =
    {
        face/2 == 0 ? (face%2 ? 1 : -1) : 0,
        face/2 == 1 ? (face%2 ? 1 : -1) : 0,
        face/2 == 2 ? (face%2 ? 1 : -1) : 0
    };
    const int fd = face/2;
    const double H = (l == 0);
    int S[3], E[3];
    layer(lv, bo, S, E);
    for (int i = S[0]; i < E[0]; i++)
    for (int j = S[1]; j < E[1]; j++)
    for (int k = S[2]; k < E[2]; k++)
    {
        const long g = lv.lin(i + bo[0], j + bo[1], k + bo[2]);
        const long p = lv.lin(i, j, k);
        if (c.bc[face] == NEUMANN)
        {
            // delta of the boundary face (upperFaceNeighbor)
            const int pc[3] = {i, j, k};
            const int fi = pc[fd] + (face%2);
            const double delta = lv.ax[fd].dl[fi];
            f[g] = f[p] + H*bcValueAt(c, face, S, E, i, j, k)/delta;
        }
        else if (c.bc[face] == DIRICHLET)
        {
            f[g] = H*2.0*bcValueAt(c, face, S, E, i, j, k) - f[p];
        }
    }
}

/* meshLevel<...>::evaluate<nonEliminatedBcs,1>: periodic faces only. */
void evaluatePeriodicFaces(const oracle_ctx& c, Level& lv,
                           std::vector<double>& f)
{
    for (int face = 0; face < 6; face++)
    {
        if (!c.periodicFace(face)) continue;
        const int bo[3] =
        {
            face/2 == 0 ? (face%2 ? 1 : -1) : 0,
            face/2 == 1 ? (face%2 ? 1 : -1) : 0,
            face/2 == 2 ? (face%2 ? 1 : -1) : 0
        };
        periodicCopy(c, lv, f, bo);
    }
}

/* meshLevel.C:231-254 + correctUnsetBoundaryConditions meshLevel.C:300-326:
 * unset edge/corner ghosts copy their interior neighbour, then patch faces,
 * then periodic faces/edges/corners. */
void correctBoundaryConditions(const oracle_ctx& c, int l, Level& lv,
                               std::vector<double>& f)
{
    for (int a = -1; a <= 1; a++)
    for (int b = -1; b <= 1; b++)
    for (int d = -1; d <= 1; d++)
    {
        const int bo[3] = {a, b, d};
        if (a == 0 && b == 0 && d == 0) continue;
        if (hasBoundary(c, bo)) continue;
        int S[3], E[3];
        layer(lv, bo, S, E);
        for (int i = S[0]; i < E[0]; i++)
        for (int j = S[1]; j < E[1]; j++)
        for (int k = S[2]; k < E[2]; k++)
            f[lv.lin(i + a, j + b, k + d)] = f[lv.lin(i, j, k)];
    }
    for (int face = 0; face < 6; face++)
        if (!c.periodicFace(face))
            patchEvaluate(c, l, lv, f, face);
    // parallel boundaries by increasing offset degree
    for (int deg = 1; deg <= 3; deg++)
    for (int a = -1; a <= 1; a++)
    for (int b = -1; b <= 1; b++)
    for (int d = -1; d <= 1; d++)
    {
        const int bo[3] = {a, b, d};
        if (std::abs(a) + std::abs(b) + std::abs(d) != deg) continue;
        bool allPeriodic = true;
        for (int q = 0; q < 3; q++)
            if (bo[q] != 0 && !c.periodicFace(2*q + (bo[q] > 0)))
                allPeriodic = false;
        if (allPeriodic)
            periodicCopy(c, lv, f, bo);
    }
}

/* Region [S,E) of rbgsSmoother.C:46-48 / linearSystemResidual.C:48-50:
 * interior padded by one cell on faces that are not eliminated. */
void region(const oracle_ctx& c, const Level& lv, int padding,
            int S[3], int E[3])
{
    for (int d = 0; d < 3; d++)
    {
        const int padLo = c.eliminated ? c.periodicFace(2*d) : 1;
        const int padHi = c.eliminated ? c.periodicFace(2*d+1) : 1;
        S[d] = 1 + padding*padLo;
        E[d] = lv.N[d] + 1 - padding*padHi;
    }
}

/* One region's four red-black phases, rbgsSmoother.C:65-183, with the
 * reference's counter-relative parity (utils.H:736-745). */
void rbgsRegion(Level& lv, const int S[3], const int E[3])
{
    const int M[3] = {E[0]-S[0], E[1]-S[1], E[2]-S[2]};
    if (M[0] <= 0 || M[1] <= 0 || M[2] <= 0) return;
    double* __restrict__ x = lv.x.data();
    const double* __restrict__ b = lv.b.data();
    const double* __restrict__ A = lv.A.data();
    const long Si = long(lv.P[1])*lv.P[2];
    const long Sj = lv.P[2];
    const long Jk = Sj - M[2];
    const long Jj = Si - long(M[1])*Sj;

    auto upd = [&](long c)
    {
        const double* a = A + 7*c;
        const double P =
            a[1]*x[c - Si]
          + a[2]*x[c + Si]
          + a[3]*x[c - Sj]
          + a[4]*x[c + Sj]
          + a[5]*x[c - 1]
          + a[6]*x[c + 1];
        x[c] = (b[c] - P)/a[0];
    };

    (void)Jj;
    // the cells of one phase read only cells of the other parity, so the
    // planes of a phase may be updated in any order (same bits)
    const bool par = long(M[0])*M[1]*M[2] > (1L << 15);
    for (int phase = 0; phase < 4; phase++)
    {
        const bool fwd = phase < 2;
        const int want = phase % 2;          // 0 even, 1 odd
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 0; i < M[0]; i++)
        {
            long c = fwd ? lv.lin(S[0] + i, S[1], S[2])
                         : lv.lin(E[0] - 1 - i, E[1] - 1, E[2] - 1);
            for (int j = 0; j < M[1]; j++)
            {
                // first k with (i+j+k)%2 == want
                int k0 = ((i + j) % 2 == want) ? 0 : 1;
                for (int k = k0; k < M[2]; k += 2)
                    upd(fwd ? c + k : c - k);
                c += fwd ? (M[2] + Jk) : -(M[2] + Jk);
            }
        }
    }
}

/* jacSmoother.C:18-127, omega 0.8 */
void jacRegion(Level& lv, const int S[3], const int E[3])
{
    const int M[3] = {E[0]-S[0], E[1]-S[1], E[2]-S[2]};
    if (M[0] <= 0 || M[1] <= 0 || M[2] <= 0) return;
    const double omega = 0.8;
    std::vector<double> y(size_t(M[0])*M[1]*M[2]);
    const long Si = long(lv.P[1])*lv.P[2];
    const long Sj = lv.P[2];
    size_t e = 0;
    for (int i = 0; i < M[0]; i++)
    for (int j = 0; j < M[1]; j++)
    for (int k = 0; k < M[2]; k++)
    {
        const long c = lv.lin(S[0]+i, S[1]+j, S[2]+k);
        const double* a = lv.A.data() + 7*c;
        const double* x = lv.x.data();
        const double P =
            a[1]*x[c - Si]
          + a[2]*x[c + Si]
          + a[3]*x[c - Sj]
          + a[4]*x[c + Sj]
          + a[5]*x[c - 1]
          + a[6]*x[c + 1];
        y[e++] = (1.0 - omega)*x[c] + omega*(lv.b[c] - P)/a[0];
    }
    e = 0;
    for (int i = 0; i < M[0]; i++)
    for (int j = 0; j < M[1]; j++)
    for (int k = 0; k < M[2]; k++)
        lv.x[lv.lin(S[0]+i, S[1]+j, S[2]+k)] = y[e++];
}

/* lexgsSmoother.C: forward then backward lexicographic sweep. */
void lexgsRegion(Level& lv, const int S[3], const int E[3])
{
    const int M[3] = {E[0]-S[0], E[1]-S[1], E[2]-S[2]};
    if (M[0] <= 0 || M[1] <= 0 || M[2] <= 0) return;
    const long Si = long(lv.P[1])*lv.P[2];
    const long Sj = lv.P[2];
    auto upd = [&](long c)
    {
        const double* a = lv.A.data() + 7*c;
        double* x = lv.x.data();
        const double P =
            a[1]*x[c - Si]
          + a[2]*x[c + Si]
          + a[3]*x[c - Sj]
          + a[4]*x[c + Sj]
          + a[5]*x[c - 1]
          + a[6]*x[c + 1];
        x[c] = (lv.b[c] - P)/a[0];
    };
    for (int i = 0; i < M[0]; i++)
    for (int j = 0; j < M[1]; j++)
    for (int k = 0; k < M[2]; k++)
        upd(lv.lin(S[0]+i, S[1]+j, S[2]+k));
    for (int i = M[0]-1; i >= 0; i--)
    for (int j = M[1]-1; j >= 0; j--)
    for (int k = M[2]-1; k >= 0; k--)
        upd(lv.lin(S[0]+i, S[1]+j, S[2]+k));
}

void sweepRegion(const oracle_ctx& c, Level& lv, const int S[3],
                 const int E[3])
{
    if (c.smoother == ORACLE_RBGS) rbgsRegion(lv, S, E);
    else if (c.smoother == ORACLE_JAC) jacRegion(lv, S, E);
    else lexgsRegion(lv, S, E);
}

/* rbgsSmoother.C:376-437 (and the identical structure of the jac/lexgs
 * Smooth): interior sweep, periodic halo fill, width-1 boundary slabs. */
void smooth(oracle_ctx& c, int l, int sweeps)
{
    Level& lv = c.L[l];
    for (int s = 0; s < sweeps; s++)
    {
        int S[3], E[3];
        region(c, lv, 1, S, E);
        sweepRegion(c, lv, S, E);
        if (!c.eliminated)
        {
            for (int face = 0; face < 6; face++)
                if (!c.periodicFace(face))
                    patchEvaluate(c, l, lv, lv.x, face);
        }
        evaluatePeriodicFaces(c, lv, lv.x);
        // SweepBoundary rbgsSmoother.C:186-374
        for (int f = 0; f < 6; f++)
        {
            const bool pad = c.eliminated ? c.periodicFace(f) : true;
            if (!pad) continue;
            const int fd = f/2;
            int Sb[3], Eb[3];
            for (int d = 0; d < 3; d++) { Sb[d] = 1; Eb[d] = lv.N[d] + 1; }
            if (f % 2 == 0) Eb[fd] = Sb[fd] + 1;
            else Sb[fd] = Eb[fd] - 1;
            sweepRegion(c, lv, Sb, Eb);
        }
    }
}

void residualRegion(Level& lv, const int S[3], const int E[3])
{
    const long Si = long(lv.P[1])*lv.P[2];
    const long Sj = lv.P[2];
    const bool par = long(E[0] - S[0])*(E[1] - S[1])*(E[2] - S[2]) > (1L << 15);
    #pragma omp parallel for schedule(static) if (par)
    for (int i = S[0]; i < E[0]; i++)
    for (int j = S[1]; j < E[1]; j++)
    {
        long c = lv.lin(i, j, S[2]);
        for (int k = S[2]; k < E[2]; k++, c++)
        {
            const double* a = lv.A.data() + 7*c;
            const double* x = lv.x.data();
            lv.r[c] =
                lv.b[c]
              - (
                    a[0]*x[c]
                  + a[1]*x[c - Si]
                  + a[2]*x[c + Si]
                  + a[3]*x[c - Sj]
                  + a[4]*x[c + Sj]
                  + a[5]*x[c - 1]
                  + a[6]*x[c + 1]
                );
        }
    }
}

/* linearSystemResidual.C:280-332 */
void residual(oracle_ctx& c, int l, bool corrBcs)
{
    Level& lv = c.L[l];
    if (corrBcs)
    {
        int S[3], E[3];
        region(c, lv, 1, S, E);
        residualRegion(lv, S, E);
        if (!c.eliminated)
            for (int face = 0; face < 6; face++)
                if (!c.periodicFace(face))
                    patchEvaluate(c, l, lv, lv.x, face);
        evaluatePeriodicFaces(c, lv, lv.x);
        for (int f = 0; f < 6; f++)
        {
            const bool pad = c.eliminated ? c.periodicFace(f) : true;
            if (!pad) continue;
            const int fd = f/2;
            int Sb[3], Eb[3];
            for (int d = 0; d < 3; d++) { Sb[d] = 1; Eb[d] = lv.N[d] + 1; }
            if (f % 2 == 0) Eb[fd] = Sb[fd] + 1;
            else Sb[fd] = Eb[fd] - 1;
            residualRegion(lv, Sb, Eb);
        }
    }
    else
    {
        int S[3], E[3];
        region(c, lv, 0, S, E);
        residualRegion(lv, S, E);
    }
}

/* linearRestrictionScheme.C:15-80 for coarse level lc (>= 1). */
void setRestrictionWeights(Level& fine, Level& coarse)
{
    coarse.rw.assign(coarse.size()*8, 0.0);
    for (int i = 1; i <= coarse.N[0]; i++)
    for (int j = 1; j <= coarse.N[1]; j++)
    for (int k = 1; k <= coarse.N[2]; k++)
    {
        const int I[3] = {i, j, k};
        double u[3];
        for (int d = 0; d < 3; d++)
        {
            // fine padded index of vertex a=0 is (I-1)*R + 1
            const int f0 = (I[d] - 1)*coarse.R[d] + 1;
            const double Px = coarse.ax[d].cc[I[d]] - fine.ax[d].cc[f0];
            const double Dx = fine.ax[d].cc[f0 + 1] - fine.ax[d].cc[f0];
            u[d] = round12((Px*Dx)/(Dx*Dx));
        }
        double* w = &coarse.rw[coarse.lin(i, j, k)*8];
        w[0] = (1.0-u[0]) * (1.0-u[1]) * (1.0-u[2]);
        w[1] = (    u[0]) * (1.0-u[1]) * (1.0-u[2]);
        w[2] = (1.0-u[0]) * (    u[1]) * (1.0-u[2]);
        w[3] = (    u[0]) * (    u[1]) * (1.0-u[2]);
        w[4] = (1.0-u[0]) * (1.0-u[1]) * (    u[2]);
        w[5] = (    u[0]) * (1.0-u[1]) * (    u[2]);
        w[6] = (1.0-u[0]) * (    u[1]) * (    u[2]);
        w[7] = (    u[0]) * (    u[1]) * (    u[2]);
        for (int q = 0; q < 8; q++)
            if (w[q] <= 1e-12) w[q] = 0.0;
    }
}

/* linearRestrictionScheme.C:103-160 with scale=true (MG.C:77). */
void restrictScaled(oracle_ctx& c, int l)
{
    Level& fine = c.L[l];
    Level& coarse = c.L[l+1];
    const int* R = coarse.R;
    const int ox = R[0] > 1, oy = R[1] > 1, oz = R[2] > 1;
    const bool par = long(coarse.N[0])*coarse.N[1]*coarse.N[2] > (1L << 12);
    #pragma omp parallel for schedule(static) if (par)
    for (int i = 1; i <= coarse.N[0]; i++)
    for (int j = 1; j <= coarse.N[1]; j++)
    for (int k = 1; k <= coarse.N[2]; k++)
    {
        const double* w = &coarse.rw[coarse.lin(i, j, k)*8];
        double v = 0.0;
        int q = 0;
        for (int cz = 0; cz < 2; cz++)
        for (int cy = 0; cy < 2; cy++)
        for (int cx = 0; cx < 2; cx++)
        {
            if (w[q++] != 0.0)
            {
                const int fi = (i-1)*R[0] + cx*ox + 1;
                const int fj = (j-1)*R[1] + cy*oy + 1;
                const int fk = (k-1)*R[2] + cz*oz + 1;
                const double cvf =
                    (fine.ax[0].w[fi]*fine.ax[1].w[fj])*fine.ax[2].w[fk];
                const double icvf = 1.0/cvf;
                v += w[q-1]*fine.r[fine.lin(fi, fj, fk)]*icvf;
            }
        }
        const double cvc =
            (coarse.ax[0].w[i]*coarse.ax[1].w[j])*coarse.ax[2].w[k];
        v *= cvc;
        coarse.b[coarse.lin(i, j, k)] = v;
    }
}

/* linearProlongationScheme.C:12-110 for fine level l. */
void setProlongationWeights(Level& fine, Level& coarse)
{
    const size_t n = size_t(fine.N[0])*fine.N[1]*fine.N[2];
    fine.pw.assign(n*8, 0.0);
    fine.pidx.assign(n*8, 0);
    size_t c = 0;
    for (int i = 0; i < fine.N[0]; i++)
    for (int j = 0; j < fine.N[1]; j++)
    for (int k = 0; k < fine.N[2]; k++)
    {
        const int ijk[3] = {i, j, k};
        int base[3], off[3];
        double u[3];
        bool inside = true;
        for (int d = 0; d < 3; d++)
        {
            off[d] = (ijk[d] % 2) ? 1 : -1;
            base[d] = ijk[d]/coarse.R[d];
            // padded indices
            const double xf = fine.ax[d].cc[ijk[d] + 1];
            const double x0 = coarse.ax[d].cc[base[d] + 1];
            const double x1 = coarse.ax[d].cc[base[d] + off[d] + 1];
            const double Px = xf - x0;
            const double Dx = x1 - x0;
            u[d] = round12((Px*Dx)/(Dx*Dx));
            if (u[d] < 0 || u[d] >= 1) inside = false;
        }
        if (!inside) u[0] = u[1] = u[2] = 1.0/2.0;
        int v = 0;
        for (int cz = 0; cz < 2; cz++)
        for (int cy = 0; cy < 2; cy++)
        for (int cx = 0; cx < 2; cx++)
        {
            const int abc[3] = {cx, cy, cz};
            fine.pidx[c*8 + v] = coarse.lin(base[0] + cx*off[0] + 1,
                                            base[1] + cy*off[1] + 1,
                                            base[2] + cz*off[2] + 1);
            double g[3];
            for (int d = 0; d < 3; d++)
                g[d] = (1.0 - abc[d] + u[d]*(2*abc[d] - 1));
            fine.pw[c*8 + v] = g[0]*g[1]*g[2];
            v++;
        }
        c++;
    }
}

/* linearProlongationScheme.C:112-147 with plusEqOp (MG.C:96). */
void prolongAdd(oracle_ctx& c, int l)
{
    Level& fine = c.L[l];
    Level& coarse = c.L[l+1];
    correctBoundaryConditions(c, l + 1, coarse, coarse.x);
    const bool par = long(fine.N[0])*fine.N[1]*fine.N[2] > (1L << 15);
    #pragma omp parallel for schedule(static) if (par)
    for (int i = 1; i <= fine.N[0]; i++)
    for (int j = 1; j <= fine.N[1]; j++)
    for (int k = 1; k <= fine.N[2]; k++)
    {
        const size_t q = (size_t(i - 1)*fine.N[1] + (j - 1))*fine.N[2] + (k - 1);
        double value = 0.0;
        for (int v = 0; v < 8; v++)
            if (fine.pw[q*8 + v] != 0.0)
                value += fine.pw[q*8 + v]*coarse.x[fine.pidx[q*8 + v]];
        double& xf = fine.x[fine.lin(i, j, k)];
        xf = xf + value;
    }
}

/* areaWeightedFaceRestrictionScheme.C:19-56 */
void restrictLambda(oracle_ctx& c, int l)
{
    Level& fine = c.L[l];
    Level& coarse = c.L[l+1];
    const int* R = coarse.R;
    for (int fd = 0; fd < 3; fd++)
    {
        coarse.lam[fd].assign(coarse.size(), 0.0);
        const int lim[3] =
        {
            coarse.N[0] + (fd == 0), coarse.N[1] + (fd == 1),
            coarse.N[2] + (fd == 2)
        };
        for (int i = 0; i < lim[0]; i++)
        for (int j = 0; j < lim[1]; j++)
        for (int k = 0; k < lim[2]; k++)
        {
            double area = 0.0, value = 0.0;
            for (int ox = 0; ox < (fd == 0 ? 1 : R[0]); ox++)
            for (int oy = 0; oy < (fd == 1 ? 1 : R[1]); oy++)
            for (int oz = 0; oz < (fd == 2 ? 1 : R[2]); oz++)
            {
                const int fi = i*R[0] + ox + 1;
                const int fj = j*R[1] + oy + 1;
                const int fk = k*R[2] + oz + 1;
                const double a = fd == 0 ? fine.ax[1].w[fj]*fine.ax[2].w[fk]
                               : fd == 1 ? fine.ax[0].w[fi]*fine.ax[2].w[fk]
                               : fine.ax[0].w[fi]*fine.ax[1].w[fj];
                area += a;
                value += a*fine.lam[fd][fine.lin(fi, fj, fk)];
            }
            coarse.lam[fd][coarse.lin(i + 1, j + 1, k + 1)] = value/area;
        }
    }
}

/* linearGaussLaplacianScheme.C:22-102 on one level. */
void assembleLevel(Level& lv, bool withLambda)
{
    lv.A.assign(lv.size()*7, 0.0);
    for (int fd = 0; fd < 3; fd++)
    {
        const int lim[3] =
        {
            lv.N[0] + (fd == 0), lv.N[1] + (fd == 1), lv.N[2] + (fd == 2)
        };
        for (int i = 1; i <= lim[0]; i++)
        for (int j = 1; j <= lim[1]; j++)
        for (int k = 1; k <= lim[2]; k++)
        {
            const int ijk[3] = {i, j, k};
            const double fa =
                fd == 0 ? lv.ax[1].w[j]*lv.ax[2].w[k]
              : fd == 1 ? lv.ax[0].w[i]*lv.ax[2].w[k]
              : lv.ax[0].w[i]*lv.ax[1].w[j];
            const double delta = lv.ax[fd].dl[ijk[fd]];
            double value = fa*delta;
            if (withLambda)
                value = fa*delta*lv.lam[fd][lv.lin(i, j, k)];
            const long P = lv.lin(i, j, k);
            const long Nn = lv.lin(i - (fd == 0), j - (fd == 1), k - (fd == 2));
            lv.A[7*P + fd*2 + 1] = value;
            lv.A[7*Nn + fd*2 + 2] = value;
            lv.A[7*P] -= value;
            lv.A[7*Nn] -= value;
        }
    }
}

/* Global cell number of padded (i,j,k) per fvMeshMetrics.C:509-557: interior
 * cells lexicographic (k fastest); face ghosts across periodic faces carry
 * the wrapped number (parallel-boundary correction), all others -1. */
long globalNumber(const oracle_ctx& c, const Level& lv, int i, int j, int k)
{
    int q[3] = {i - 1, j - 1, k - 1};
    int nout = 0;
    for (int d = 0; d < 3; d++)
    {
        if (q[d] < 0 || q[d] >= lv.N[d])
        {
            nout++;
            const int face = 2*d + (q[d] >= lv.N[d]);
            if (!c.periodicFace(face)) return -1;
            q[d] = (q[d] + lv.N[d]) % lv.N[d];
        }
    }
    (void)nout;
    return (long(q[0])*lv.N[1] + q[1])*lv.N[2] + q[2];
}

/* solver.C:200-231 on level l */
void setSingularityConstraint(oracle_ctx& c, int l)
{
    if (!c.singular()) return;
    Level& lv = c.L[l];
    double* a0 = &lv.A[7*lv.lin(1, 1, 1)];
    a0[0] = 1.0;
    for (int s = 1; s < 7; s++) a0[s] = 0.0;
    for (int i = 1; i <= lv.N[0]; i++)
    for (int j = 1; j <= lv.N[1]; j++)
    for (int k = 1; k <= lv.N[2]; k++)
    {
        double* a = &lv.A[7*lv.lin(i, j, k)];
        for (int s = 1; s < 7; s++)
            if (a[s] != 0)
                if (!globalNumber(c, lv, i + OFF[s][0], j + OFF[s][1],
                                  k + OFF[s][2]))
                    a[s] = 0;
    }
}

/* linearSystem.C:325-341 with Neumann/Dirichlet eliminateGhosts
 * (NeumannBoundaryCondition.C:19-56, DirichletBoundaryCondition.C:18-48). */
void eliminateGhosts(oracle_ctx& c)
{
    for (int l = 0; l < c.nLevels(); l++)
    {
        Level& lv = c.L[l];
        for (int face = 0; face < 6; face++)
        {
            if (c.bc[face] == PERIODIC) continue;
            const int fd = face/2;
            const int bo[3] =
            {
                fd == 0 ? (face%2 ? 1 : -1) : 0,
                fd == 1 ? (face%2 ? 1 : -1) : 0,
                fd == 2 ? (face%2 ? 1 : -1) : 0
            };
            int S[3], E[3];
            layer(lv, bo, S, E);
            for (int i = S[0]; i < E[0]; i++)
            for (int j = S[1]; j < E[1]; j++)
            for (int k = S[2]; k < E[2]; k++)
            {
                const long p = lv.lin(i, j, k);
                double* a = &lv.A[7*p];
                const double g = a[face + 1];
                if (c.bc[face] == NEUMANN)
                {
                    a[0] += g;
                    a[face + 1] = 0.0;
                    if (l == 0)
                    {
                        const int pc[3] = {i, j, k};
                        const double delta =
                            lv.ax[fd].dl[pc[fd] + (face%2)];
                        lv.b[p] -= g*bcValueAt(c, face, S, E, i, j, k)
                                 / delta;
                    }
                }
                else
                {
                    a[0] -= g;
                    a[face + 1] = 0.0;
                    if (l == 0)
                        lv.b[p] -= g*2.0*bcValueAt(c, face, S, E, i, j, k);
                }
            }
        }
    }
    c.eliminated = true;
}

/* Restated Crout LU with implicit (scaled) partial pivoting. */
void luDecompose(std::vector<double>& m, std::vector<int>& piv, int n)
{
    std::vector<double> vv(n);
    piv.assign(n, 0);
    for (int i = 0; i < n; i++)
    {
        double big = 0.0;
        for (int j = 0; j < n; j++)
            big = std::max(big, std::fabs(m[i*n + j]));
        vv[i] = big > 0.0 ? 1.0/big : 0.0;
    }
    for (int j = 0; j < n; j++)
    {
        for (int i = 0; i < j; i++)
        {
            double sum = m[i*n + j];
            for (int k = 0; k < i; k++)
                sum -= m[i*n + k]*m[k*n + j];
            m[i*n + j] = sum;
        }
        int imax = j;
        double big = 0.0;
        for (int i = j; i < n; i++)
        {
            double sum = m[i*n + j];
            for (int k = 0; k < j; k++)
                sum -= m[i*n + k]*m[k*n + j];
            m[i*n + j] = sum;
            const double t = vv[i]*std::fabs(sum);
            if (t >= big) { big = t; imax = i; }
        }
        piv[j] = imax;
        if (j != imax)
        {
            for (int k = 0; k < n; k++)
                std::swap(m[imax*n + k], m[j*n + k]);
            vv[imax] = vv[j];
        }
        if (m[j*n + j] == 0.0) m[j*n + j] = 1e-15;
        if (j != n - 1)
        {
            const double r = 1.0/m[j*n + j];
            for (int i = j + 1; i < n; i++)
                m[i*n + j] *= r;
        }
    }
}

void luBacksubstitute(const std::vector<double>& m, const std::vector<int>& piv,
                      std::vector<double>& b, int n)
{
    for (int i = 0; i < n; i++)
    {
        const int ip = piv[i];
        double sum = b[ip];
        b[ip] = b[i];
        for (int j = 0; j < i; j++)
            sum -= m[i*n + j]*b[j];
        b[i] = sum;
    }
    for (int i = n - 1; i >= 0; i--)
    {
        double sum = b[i];
        for (int j = i + 1; j < n; j++)
            sum -= m[i*n + j]*b[j];
        b[i] = sum/m[i*n + i];
    }
}

/* APLU::prepare (APLU.C:26-89) via linearSystemAggregation::rowCoeffs and
 * compressedRowFormat (linearSystemAggregation.C:236-330, 434-517). */
void prepareCoarse(oracle_ctx& c)
{
    const int l = c.coarseLevel();
    Level& lv = c.L[l];
    const int n = lv.N[0]*lv.N[1]*lv.N[2];
    c.nc = n;
    c.D.assign(size_t(n)*n, 0.0);
    for (int i = 1; i <= lv.N[0]; i++)
    for (int j = 1; j <= lv.N[1]; j++)
    for (int k = 1; k <= lv.N[2]; k++)
    {
        const long row = globalNumber(c, lv, i, j, k);
        double a[7];
        long col[7];
        for (int s = 0; s < 7; s++)
        {
            a[s] = lv.A[7*lv.lin(i, j, k) + s];
            col[s] = globalNumber(c, lv, i + OFF[s][0], j + OFF[s][1],
                                  k + OFF[s][2]);
            if (col[s] < 0 || col[s] >= n) a[s] = 0;
        }
        for (int s = 1; s < 7; s++)
        {
            int t = 0;
            while (col[t] != col[s]) t++;
            if (t < s) { a[t] += a[s]; a[s] = 0; }
        }
        for (int s = 0; s < 7; s++)
            if (a[s] != 0)
                c.D[size_t(row)*n + col[s]] = a[s];
    }
    c.LU = c.D;
    luDecompose(c.LU, c.piv, n);
    c.prepared = true;
}

/* APLU::solve (APLU.C:91-128) with rhsSource (linearSystemAggregation.C:
 * 336-430). */
void coarseSolve(oracle_ctx& c)
{
    if (!c.prepared) prepareCoarse(c);
    const int l = c.coarseLevel();
    Level& lv = c.L[l];
    const int n = c.nc;
    std::vector<double> rhs(n);
    int q = 0;
    for (int i = 1; i <= lv.N[0]; i++)
    for (int j = 1; j <= lv.N[1]; j++)
    for (int k = 1; k <= lv.N[2]; k++)
    {
        const long p = lv.lin(i, j, k);
        rhs[q] = lv.b[p];
        for (int s = 0; s < 7; s++)
        {
            const int ii = i + OFF[s][0], jj = j + OFF[s][1],
                      kk = k + OFF[s][2];
            const long g = globalNumber(c, lv, ii, jj, kk);
            if (g < 0 || g >= n)
                rhs[q] -= lv.A[7*p + s]*lv.x[lv.lin(ii, jj, kk)];
        }
        q++;
    }
    luBacksubstitute(c.LU, c.piv, rhs, n);
    q = 0;
    for (int i = 1; i <= lv.N[0]; i++)
    for (int j = 1; j <= lv.N[1]; j++)
    for (int k = 1; k <= lv.N[2]; k++)
        lv.x[lv.lin(i, j, k)] = rhs[q++];
    correctBoundaryConditions(c, l, lv, lv.x);
}

int levelReps(const oracle_ctx& c, int l, const std::vector<int>& visits)
{
    if (c.cycle == ORACLE_VCYCLE) return 1;
    if (c.cycle == ORACLE_WCYCLE) return 1 + (l > 0);
    return 1 + (l > 0 && visits[l] == 0);
}

/* MG.C:19-129 */
void cycleLevel(oracle_ctx& c, std::vector<int>& visits, int l)
{
    Level& lv = c.L[l];
    const int lc = c.coarseLevel();
    if (c.singular() && l == lc)
        lv.b[lv.lin(1, 1, 1)] = 0.0;
    if (l > 0)
        std::fill(lv.x.begin(), lv.x.end(), 0.0);
    if (c.nPre > 0)
    {
        smooth(c, l, c.nPre);
        c.cnt[1][l]++;
    }
    if (l < lc)
    {
        for (int rep = 0; rep < levelReps(c, l, visits); rep++)
        {
            if (rep > 0 || l > 0)
            {
                residual(c, l, true);
                c.cnt[2][l]++;
            }
            restrictScaled(c, l);
            c.cnt[3][l]++;
            cycleLevel(c, visits, l + 1);
            prolongAdd(c, l);
            c.cnt[4][l]++;
            if (c.nPost > 0)
            {
                smooth(c, l, c.nPost);
                c.cnt[1][l]++;
            }
        }
    }
    else
    {
        if (c.coarseMode == ORACLE_COARSE_DIRECT)
        {
            coarseSolve(c);
            c.cnt[5][l]++;
        }
        else
        {
            smooth(c, l, std::max(c.nPost, 2));
            c.cnt[1][l]++;
        }
    }
    visits[l]++;
    c.cnt[0][l]++;
}

void resetCounts(oracle_ctx& c)
{
    for (auto& v : c.cnt) v.assign(c.nLevels(), 0);
}

double sumSq(const Level& lv, const std::vector<double>& f)
{
    double s = 0.0;
    for (int i = 1; i <= lv.N[0]; i++)
    for (int j = 1; j <= lv.N[1]; j++)
    for (int k = 1; k <= lv.N[2]; k++)
    {
        const double v = f[lv.lin(i, j, k)];
        s += v*v;
    }
    return s;
}

/* solver.C:13-61 */
double normFactor(const oracle_ctx& c)
{
    const Level& lv = c.L[0];
    const double n = double(lv.N[0])*lv.N[1]*lv.N[2];
    double sum = 0.0;
    for (int i = 1; i <= lv.N[0]; i++)
    for (int j = 1; j <= lv.N[1]; j++)
    for (int k = 1; k <= lv.N[2]; k++)
        sum += lv.x[lv.lin(i, j, k)];
    const double y = sum/n;
    double f = 0.0;
    for (int i = 1; i <= lv.N[0]; i++)
    for (int j = 1; j <= lv.N[1]; j++)
    for (int k = 1; k <= lv.N[2]; k++)
    {
        const long p = lv.lin(i, j, k);
        const double* a = &lv.A[7*p];
        double rs = a[0];
        for (int s = 1; s < 7; s++) rs += a[s];
        const double Ay = rs*y;
        const double t = std::fabs(lv.r[p] - lv.b[p] + Ay)
                       + std::fabs(lv.b[p] - Ay);
        f += t*t;
    }
    f = std::sqrt(f);
    return std::max(f, n*1e-12);
}

/* solver.C:63-95 */
int checkConvergence(double cur, double init, double relTol, double absTol)
{
    return (init > 0.0 && cur > absTol && cur/(init + 1e-20) > relTol) ? 0 : 1;
}

} // namespace

extern "C" {

oracle_ctx* oracle_create(const int n[3], const double* vx, const double* vy,
                          const double* vz, const int bc[6],
                          const double bc_value[6])
{
    oracle_ctx* c = new oracle_ctx;
    for (int f = 0; f < 6; f++) { c->bc[f] = bc[f]; c->bcval[f] = bc_value[f]; }
    const double* verts[3] = {vx, vy, vz};

    // mesh::generateLevels (mesh.C:19-53) with level::NandR (level.C:11-53)
    int N[3] = {n[0], n[1], n[2]};
    int R[3] = {0, 0, 0};
    while (true)
    {
        Level lv;
        for (int d = 0; d < 3; d++)
        {
            lv.N[d] = N[d];
            lv.P[d] = N[d] + 2;
            lv.R[d] = R[d];
            lv.ax[d].N = N[d];
            lv.ax[d].R = R[d];
            std::vector<double> v(N[d] + 1);
            if (c->L.empty())
                for (int q = 0; q <= N[d]; q++) v[q] = verts[d][q];
            else
            {
                const Axis& pa = c->L.back().ax[d];
                for (int q = 0; q <= N[d]; q++) v[q] = pa.v[q*R[d] + 1];
            }
            buildAxis(lv.ax[d], v, c->periodicFace(2*d));
            axisMetrics(lv.ax[d], d);
        }
        lv.x.assign(lv.size(), 0.0);
        lv.b.assign(lv.size(), 0.0);
        lv.r.assign(lv.size(), 0.0);
        lv.A.assign(lv.size()*7, 0.0);
        c->L.push_back(std::move(lv));

        // coarsen (utils.H:1154-1171)
        int Nc[3];
        bool coarsenable = false;
        for (int d = 0; d < 3; d++)
        {
            Nc[d] = N[d] <= 3 ? N[d] : N[d]/2;
            if (Nc[d] != N[d]) coarsenable = true;
        }
        if (!coarsenable) break;
        for (int d = 0; d < 3; d++) { R[d] = N[d]/Nc[d]; N[d] = Nc[d]; }
    }
    for (size_t l = 0; l + 1 < c->L.size(); l++)
    {
        setRestrictionWeights(c->L[l], c->L[l+1]);
        setProlongationWeights(c->L[l], c->L[l+1]);
    }
    return c;
}

void oracle_destroy(oracle_ctx* c) { delete c; }

/* Threads of the per-cell loops (OpenMP; results do not depend on it). */
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_get_threads(void) { return omp_get_max_threads(); }

void oracle_set_opts(oracle_ctx* c, double rel_tol, double tolerance,
                     int min_iter, int max_iter, int n_sweeps_pre,
                     int n_sweeps_post, int cycle, int coarse_mode,
                     int coarse_level, int smoother)
{
    c->relTol = rel_tol;
    c->tolerance = tolerance;
    c->minIter = min_iter;
    c->maxIter = max_iter;
    c->nPre = n_sweeps_pre;
    c->nPost = n_sweeps_post;
    c->cycle = cycle;
    c->coarseMode = coarse_mode;
    c->coarseLevel_ = coarse_level;
    c->smoother = smoother;
    c->prepared = false;
}

int oracle_num_levels(const oracle_ctx* c) { return c->nLevels(); }

void oracle_level_shape(const oracle_ctx* c, int l, int n[3])
{
    for (int d = 0; d < 3; d++) n[d] = c->L[l].N[d];
}

void oracle_level_ratio(const oracle_ctx* c, int l, int r[3])
{
    for (int d = 0; d < 3; d++) r[d] = c->L[l].R[d];
}

void oracle_level_centres(const oracle_ctx* c, int l, int axis, double* out)
{
    const Axis& a = c->L[l].ax[axis];
    std::copy(a.cc.begin(), a.cc.end(), out);
}

void oracle_level_widths(const oracle_ctx* c, int l, int axis, double* out)
{
    const Axis& a = c->L[l].ax[axis];
    std::copy(a.w.begin(), a.w.end(), out);
}

void oracle_assemble_laplacian(oracle_ctx* c, const double* lam_x,
                               const double* lam_y, const double* lam_z)
{
    oracle_assemble_laplacian_ddt(c, lam_x, lam_y, lam_z, 0.0, 1.0);
}

/* sys -= im::ddt(x)/dtFrac (defaultPoissonSolver.C:91-92) after the
 * Laplacian: the Euler matrix is cv/deltaT on every level
 * (EulerDdtScheme.C:50-54), divided by dtFrac (linearSystemI.H:392-410),
 * and subtracted from the stencil centre (linearSystem.C:694-706).  Volumes
 * are (dx*dy)*dz; only interior rows are solved, so ghost rows keep the
 * Laplacian alone. */
void oracle_assemble_laplacian_ddt(oracle_ctx* c, const double* lam_x,
                                   const double* lam_y, const double* lam_z,
                                   double delta_t, double dt_frac)
{
    const bool withLambda = lam_x != nullptr;
    if (withLambda)
    {
        const double* lam[3] = {lam_x, lam_y, lam_z};
        for (int fd = 0; fd < 3; fd++)
            c->L[0].lam[fd].assign(lam[fd], lam[fd] + c->L[0].size());
        for (int l = 0; l + 1 < c->nLevels(); l++)
            restrictLambda(*c, l);
    }
    for (auto& lv : c->L)
    {
        assembleLevel(lv, withLambda);
        std::fill(lv.b.begin(), lv.b.end(), 0.0);
        if (delta_t == 0.0) continue;
        for (int i = 1; i <= lv.N[0]; i++)
        for (int j = 1; j <= lv.N[1]; j++)
        for (int k = 1; k <= lv.N[2]; k++)
        {
            const double cv = (lv.ax[0].w[i]*lv.ax[1].w[j])*lv.ax[2].w[k];
            const double a = cv/delta_t;
            lv.A[7*lv.lin(i, j, k)] -= a/dt_frac;
        }
    }
    c->symmetric = true;
    c->eliminated = false;
    c->prepared = false;
}

void oracle_set_stencil(oracle_ctx* c, int l, const double* A7)
{
    Level& lv = c->L[l];
    std::copy(A7, A7 + lv.size()*7, lv.A.begin());
    c->eliminated = false;
    c->prepared = false;
}

void oracle_get_stencil(const oracle_ctx* c, int l, double* A7)
{
    const Level& lv = c->L[l];
    std::copy(lv.A.begin(), lv.A.end(), A7);
}

void oracle_set_symmetric(oracle_ctx* c, int s) { c->symmetric = s != 0; }
int oracle_singular(const oracle_ctx* c) { return c->singular(); }

static std::vector<double>& fieldOf(oracle_ctx* c, int which, int l)
{
    Level& lv = c->L[l];
    return which == ORACLE_X ? lv.x : which == ORACLE_B ? lv.b : lv.r;
}

void oracle_set_field(oracle_ctx* c, int which, int l, const double* v)
{
    auto& f = fieldOf(c, which, l);
    std::copy(v, v + f.size(), f.begin());
}

void oracle_get_field(const oracle_ctx* c, int which, int l, double* v)
{
    auto& f = fieldOf(const_cast<oracle_ctx*>(c), which, l);
    std::copy(f.begin(), f.end(), v);
}

void oracle_pin_and_eliminate(oracle_ctx* c)
{
    setSingularityConstraint(*c, c->coarseLevel());
    eliminateGhosts(*c);
}

void oracle_prepare_coarse(oracle_ctx* c) { prepareCoarse(*c); }
void oracle_smooth(oracle_ctx* c, int l, int sweeps) { smooth(*c, l, sweeps); }
void oracle_residual(oracle_ctx* c, int l, int corr) { residual(*c, l, corr); }
void oracle_restrict(oracle_ctx* c, int l) { restrictScaled(*c, l); }
void oracle_prolong(oracle_ctx* c, int l) { prolongAdd(*c, l); }

void oracle_correct_bcs(oracle_ctx* c, int l)
{
    correctBoundaryConditions(*c, l, c->L[l], c->L[l].x);
}

void oracle_coarse_solve(oracle_ctx* c) { coarseSolve(*c); }

void oracle_cycle(oracle_ctx* c)
{
    std::vector<int> visits(c->nLevels(), 0);
    resetCounts(*c);
    cycleLevel(*c, visits, 0);
}

void oracle_cycle_counts(const oracle_ctx* c, int what, long* out)
{
    for (int l = 0; l < c->nLevels(); l++)
        out[l] = (what >= 0 && what < 6 && !c->cnt[what].empty())
               ? c->cnt[what][l] : 0;
}

int oracle_set_boundary_values(oracle_ctx* c, int face, const double* v)
{
    if (face < 0 || face > 5) return -1;
    if (!v) { c->bcarr[face].clear(); return 0; }
    const int fd = face/2;
    long n = 1;
    for (int d = 0; d < 3; d++)
        if (d != fd) n *= c->L[0].N[d];
    c->bcarr[face].assign(v, v + n);
    return 0;
}

void oracle_correct_eliminated(oracle_ctx* c)
{
    for (int face = 0; face < 6; face++)
        if (!c->periodicFace(face))
            patchEvaluate(*c, 0, c->L[0], c->L[0].x, face);
}

double oracle_sum_sq_residual(const oracle_ctx* c)
{
    return sumSq(c->L[0], c->L[0].r);
}

void oracle_norm_factor(const oracle_ctx* c, double* nf)
{
    *nf = normFactor(*c);
}

/* MG::solve (MG.C:402-483) around MG::Solve (MG.C:132-254). */
void oracle_solve(oracle_ctx* c, int const_matrix, double stats[3])
{
    setSingularityConstraint(*c, c->coarseLevel());
    eliminateGhosts(*c);
    if (c->coarseMode == ORACLE_COARSE_DIRECT && (!const_matrix || !c->prepared))
        prepareCoarse(*c);

    Level& l0 = c->L[0];
    std::fill(l0.r.begin(), l0.r.end(), 0.0);
    residual(*c, 0, false);
    const double nf = normFactor(*c);
    const double init = std::sqrt(sumSq(l0, l0.r))/nf;
    double cur = init;
    int conv = checkConvergence(cur, init, c->relTol, c->tolerance);
    int iter = 0;
    std::vector<int> visits(c->nLevels(), 0);
    while ((iter < c->maxIter && conv == 0) || iter < c->minIter)
    {
        if (iter < c->minIter) conv = 0;
        std::fill(visits.begin(), visits.end(), 0);
        resetCounts(*c);
        cycleLevel(*c, visits, 0);
        residual(*c, 0, true);
        cur = std::sqrt(sumSq(l0, l0.r))/nf;
        conv = checkConvergence(cur, init, c->relTol, c->tolerance);
        iter++;
    }
    oracle_correct_eliminated(c);
    stats[0] = iter;
    stats[1] = init;
    stats[2] = cur;
}

/* MG::solve for a diagonal system (MG.C:404-428): ghost elimination, one
 * diagonalSmoother::Sweep (x = b/A_C, diagonalSmoother.C:13-69) and the
 * level-0 boundary update. */
void oracle_solve_diagonal(oracle_ctx* c)
{
    eliminateGhosts(*c);
    Level& lv = c->L[0];
    for (int i = 1; i <= lv.N[0]; i++)
    for (int j = 1; j <= lv.N[1]; j++)
    for (int k = 1; k <= lv.N[2]; k++)
    {
        const long p = lv.lin(i, j, k);
        lv.x[p] = lv.b[p]/lv.A[7*p];
    }
    correctBoundaryConditions(*c, 0, lv, lv.x);
}

int oracle_coarse_dense(const oracle_ctx* c, double* dense, double* lu,
                        int* pivots)
{
    if (!c->prepared) return 0;
    const int n = c->nc;
    if (dense) std::copy(c->D.begin(), c->D.end(), dense);
    if (lu) std::copy(c->LU.begin(), c->LU.end(), lu);
    if (pivots) std::copy(c->piv.begin(), c->piv.end(), pivots);
    return n;
}

} // extern "C"
