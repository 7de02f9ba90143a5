// FFT Poisson solver on one MI355X: the reference's FFTPoissonSolver
// (src/briscolaFiniteVolume/solvers/PoissonSolvers/FFTPoissonSolver/) for a
// rectilinear mesh with at least two uniform directions, solving
// div(grad x) = -b in its finite-difference form.
//
// Steps (one stream, no host synchronisation inside a solve):
//   1. right-hand side -b minus the boundary terms of the boundary cells
//      (FFTPoissonSolver.C:123-180), scattered into a complex array E[s][a][c]
//      (s the tridiagonal axis, a and c the transformed axes); a Neumann or
//      Dirichlet transformed axis is extended to 2N by even or odd
//      reflection, so that one complex FFT of the extension gives the
//      DCT-II / DST-II (FFTW REDFT10 / RODFT10) after a twiddle, and a
//      Dirichlet-Neumann or Neumann-Dirichlet axis to 4N by the quarter-wave
//      reflections, whose FFT gives the DST-IV / DCT-IV (RODFT11 / REDFT11)
//      at the odd positions;
//   2. one batched 2-D complex FFT over (a, c) (rocFFT);
//   3. per wave-number pair, the twiddle and the tridiagonal solve along s
//      with the reference's coefficients, boundary rows, pinned zero mode
//      and, for a periodic s, its cyclic two-solve variant
//      (tridiagonalSolver.C:14-475), into a compact mode array;
//   4. the modes expanded back to the extended spectrum (inverse twiddle,
//      the reflection symmetries) and one batched inverse 2-D FFT;
//   5. the real part scaled by 1/(L_a L_c) (the reference's N or 2N per
//      transformed axis) into x, then every ghost of x refreshed
//      (x.correctBoundaryConditions(), FFTPoissonSolver.C:437-440).
// Every boundary pair is taken on every axis.
#include "bmg_context.h"

#include <rocfft/rocfft.h>

#include <cmath>
#include <mutex>

namespace bmg {

namespace {

// Axis codes of FourierTransforms.C:133-215.
enum { C_DD = 1, C_NN = 2, C_DN = 3, C_ND = 4, C_PER = 5 };

int axisCode(int lo, int hi)
{
    if (lo == PER) return C_PER;
    if (lo == DIR) return hi == DIR ? C_DD : C_DN;
    return hi == DIR ? C_ND : C_NN;
}

struct FftGeom {
    int N[3], P[3];
    int s, d1, d2;          // tridiagonal axis, transformed axes
    int n1, n2, ns;         // sizes along d1, d2, s
    int L1, L2;             // transform lengths (N, 2N or 4N)
    int code1, code2, codes;
    int bc[6];
    double bcval[6];
    double hwall[6];        // width of the boundary cell of each face
    double d1sqr, d2sqr;
    double dT;              // deltaT*dtFrac of the time term; 0: none
};

struct C2 {
    double x, y;
};

__device__ __forceinline__ C2 cmul(C2 a, C2 b)
{
    return C2{a.x*b.x - a.y*b.y, a.x*b.y + a.y*b.x};
}

// e^{i theta}
__device__ __forceinline__ C2 cexpi(double th)
{
    double sn, cs;
    sincos(th, &sn, &cs);
    return C2{cs, sn};
}

__device__ __forceinline__ long refIndex(const FftGeom& g, int i, int j,
                                         int k)
{
    return ((long)(i + 1)*g.P[1] + (j + 1))*g.P[2] + (k + 1);
}

// The positions of value a of a transformed axis in its extension and their
// signs: periodic a; Neumann-Neumann (even about -1/2 and N-1/2, period 2N)
// a, 2N-1-a; Dirichlet-Dirichlet (odd about both) a, -(2N-1-a);
// Neumann-Dirichlet (DCT-IV: even about -1/2, odd about N-1/2, period 4N)
// a, -(2N-1-a), -(2N+a), 4N-1-a; Dirichlet-Neumann (DST-IV: odd about -1/2,
// even about N-1/2) a, 2N-1-a, -(2N+a), -(4N-1-a).
__device__ __forceinline__ int extension(int code, int n, int a, int p[4],
                                         double w[4])
{
    p[0] = a;
    w[0] = 1.0;
    if (code == C_PER) return 1;
    p[1] = 2*n - 1 - a;
    if (code == C_NN || code == C_DD)
    {
        w[1] = code == C_NN ? 1.0 : -1.0;
        return 2;
    }
    p[2] = 2*n + a;
    p[3] = 4*n - 1 - a;
    const bool dct = code == C_ND;
    w[1] = dct ? -1.0 : 1.0;
    w[2] = -1.0;
    w[3] = dct ? 1.0 : -1.0;
    return 4;
}

// Step 1: interior cell t of the reference layout.
__global__ void k_fft_rhs(FftGeom g, const double* b, const double* x,
                          const double* xold, C2* E)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)g.N[0]*g.N[1]*g.N[2]) return;
    int q[3];
    q[0] = int(t/((long)g.N[1]*g.N[2]));
    q[1] = int((t/g.N[2]) % g.N[1]);
    q[2] = int(t % g.N[2]);
    const long c = refIndex(g, q[0], q[1], q[2]);
    double f = b ? -b[c] : 0.0;
    // ddt: minus x.oldTime()/(deltaT*dtFrac) (FFTPoissonSolver.C:116-131)
    if (xold) f = f - xold[c]/g.dT;
    const long stride[3] = {(long)g.P[1]*g.P[2], (long)g.P[2], 1};
    for (int face = 0; face < 6; face++)
    {
        const int d = face >> 1, side = face & 1;
        if (g.bc[face] == PER) continue;
        if (q[d] != (side ? g.N[d] - 1 : 0)) continue;
        const double h = g.hwall[face];
        const double sq = h*h;
        const double xg = x[c + (side ? stride[d] : -stride[d])];
        const double xc = x[c];
        const double term = g.bc[face] == DIR ? (xg + xc)/sq : (xg - xc)/sq;
        f -= term;
    }
    const int s = q[g.s], a = q[g.d1], cc = q[g.d2];
    int p1[4], p2[4];
    double w1[4], w2[4];
    const int c1 = extension(g.code1, g.n1, a, p1, w1);
    const int c2 = extension(g.code2, g.n2, cc, p2, w2);
    const long base = (long)s*g.L1*g.L2;
    for (int u = 0; u < c1; u++)
        for (int v = 0; v < c2; v++)
            E[base + (long)p1[u]*g.L2 + p2[v]] = C2{w1[u]*w2[v]*f, 0.0};
}

// Mode k of a transformed axis: its position m in the spectrum of the
// (extended) FFT and the factor turning that FFT into the transform:
// periodic X_k = Y_k; DCT-II X_k = e^{-i pi k/2N} Y_k; DST-II X_k =
// i e^{-i pi (k+1)/2N} Y_{k+1}; DCT-IV X_k = 1/2 e^{-i phi} Y_{2k+1} and
// DST-IV X_k = i/2 e^{-i phi} Y_{2k+1} with phi = pi (2k+1)/4N (the sum of
// the four reflected copies is twice the transform).
__device__ __forceinline__ void fwdMode(int code, int n, int k, int& m,
                                        C2& w)
{
    if (code == C_PER) { m = k; w = C2{1.0, 0.0}; return; }
    const double pi = 3.14159265358979323846;
    if (code == C_NN)
    {
        m = k;
        w = cexpi(-pi*k/(2.0*n));
        return;
    }
    if (code == C_ND || code == C_DN)
    {
        m = 2*k + 1;
        const C2 e = cexpi(-pi*m/(4.0*n));
        w = code == C_ND ? C2{0.5*e.x, 0.5*e.y} : C2{-0.5*e.y, 0.5*e.x};
        return;
    }
    m = k + 1;
    const C2 e = cexpi(-pi*m/(2.0*n));
    w = C2{-e.y, e.x};   // i*e
}

// Eigenvalue of mode k (tridiagonalSolver.C:31-98).
__device__ __forceinline__ double eigen(int code, int n, int k)
{
    const double pi = 3.14159265358979323846;
    double sn;
    if (code == C_DD) sn = sin((k + 1)*pi/(2.0*n));
    else if (code == C_NN) sn = sin(k*pi/(2.0*n));
    else if (code == C_PER) sn = sin(k*pi/n);
    else sn = sin((2.0*(k + 1) - 1.0)*pi/(4.0*n));
    return -4.0*(sn*sn);
}

// Step 3: one thread per wave-number pair (k1, k2); the modes of all s are
// read from E, solved, and written to X[s][k1][k2].  Dh and V are scratch
// of the same shape (eliminated diagonal, cyclic auxiliary solution).
__global__ void k_fft_tridiag(FftGeom g, const C2* E, C2* X, double* Dh,
                              double* V, const double* h, const double* de,
                              const double* dw)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    const long nm = (long)g.n1*g.n2;
    if (t >= nm) return;
    const int k1 = int(t/g.n2), k2 = int(t % g.n2);
    int m1, m2;
    C2 w1, w2;
    fwdMode(g.code1, g.n1, k1, m1, w1);
    fwdMode(g.code2, g.n2, k2, m2, w2);
    const C2 w = cmul(w1, w2);
    const double lam1 = eigen(g.code1, g.n1, k1);
    const double lam2 = eigen(g.code2, g.n2, k2);
    const double Lm = lam1/g.d1sqr + lam2/g.d2sqr;
    const int ns = g.ns, last = ns - 1;
    const bool zero = lam1 == 0.0 && lam2 == 0.0;
    // main diagonal of row k (tridiagonalSolver.C:107-250)
    auto diag = [&](int k) -> double {
        const double east = 1.0/(de[k]*h[k]);
        const double west = 1.0/(h[k]*dw[k]);
        if (k > 0 && k < last) return Lm - east - west;
        const double pin = -1.0/(de[0]*h[0]) - 2.0/(h[0]*dw[0]);
        switch (g.codes)
        {
        case C_DD:
            return k == 0 ? Lm - east - 2.0/(h[0]*dw[0])
                          : Lm - 2.0/(de[last]*h[last]) - west;
        case C_NN:
            if (k == 0) return zero ? pin : Lm - east;
            return Lm - west;
        case C_DN:
            return k == 0 ? Lm - east - 2.0/(h[0]*dw[0]) : Lm - west;
        case C_ND:
            return k == 0 ? Lm - east
                          : Lm - 2.0/(de[last]*h[last]) - west;
        default:
            if (k == 0) return zero ? pin : Lm - east - west;
            return Lm - east - west;
        }
    };
    // ddt: every diagonal minus 1/(deltaT*dtFrac) (tridiagonalSolver.C:352,
    // :430), the pinned zero-mode row included
    auto diagT = [&](int k) -> double {
        const double d = diag(k);
        return g.dT != 0.0 ? d - 1.0/g.dT : d;
    };
    auto DU = [&](int k) { return 1.0/(h[k]*de[k]); };
    auto DL = [&](int k) { return 1.0/(h[k]*dw[k]); };
    const long plane = (long)g.L1*g.L2;
    const long off = (long)m1*g.L2 + m2;
    const bool cyc = g.codes == C_PER;
    const int start = cyc ? 1 : 0;
    // right-hand sides: the transformed modes
    for (int k = 0; k < ns; k++)
    {
        const C2 y = E[k*plane + off];
        X[k*nm + t] = cmul(w, y);
        Dh[k*nm + t] = diagT(k);
        if (cyc) V[k*nm + t] = k == 1 ? -DL(1) : (k == last ? -DU(last) : 0.0);
    }
    // forward elimination (tridiagonalSolver.C:340-346)
    for (int k = 1 + start; k < ns; k++)
    {
        const double dp = Dh[(k - 1)*nm + t];
        const double m = DL(k)/dp;
        Dh[k*nm + t] = Dh[k*nm + t] - m*DU(k - 1);
        C2 f = X[k*nm + t];
        const C2 fp = X[(k - 1)*nm + t];
        f.x = f.x - m*fp.x;
        f.y = f.y - m*fp.y;
        X[k*nm + t] = f;
        if (cyc) V[k*nm + t] = V[k*nm + t] - m*V[(k - 1)*nm + t];
    }
    // back substitution (:348-354)
    {
        const double d = Dh[last*nm + t];
        C2 p = X[last*nm + t];
        p.x = p.x/d;
        p.y = p.y/d;
        X[last*nm + t] = p;
        if (cyc) V[last*nm + t] = V[last*nm + t]/d;
    }
    for (int k = ns - 2; k >= start; k--)
    {
        const double d = Dh[k*nm + t];
        const C2 pn = X[(k + 1)*nm + t];
        C2 f = X[k*nm + t];
        f.x = (f.x - DU(k)*pn.x)/d;
        f.y = (f.y - DU(k)*pn.y)/d;
        X[k*nm + t] = f;
        if (cyc) V[k*nm + t] = (V[k*nm + t] - DU(k)*V[(k + 1)*nm + t])/d;
    }
    if (cyc)
    {
        // reconstruction (tridiagonalSolver.C:418-450)
        const C2 f0 = cmul(w, E[off]);
        const C2 uN = X[last*nm + t], u1 = X[nm + t];
        const double vN = V[last*nm + t], v1 = V[nm + t];
        const double den = (diagT(0) + DL(0)*vN) + DU(0)*v1;
        C2 p0;
        p0.x = ((f0.x - DL(0)*uN.x) - DU(0)*u1.x)/den;
        p0.y = ((f0.y - DL(0)*uN.y) - DU(0)*u1.y)/den;
        X[t] = p0;
        for (int k = 1; k < ns; k++)
        {
            C2 u = X[k*nm + t];
            const double v = V[k*nm + t];
            u.x = u.x + p0.x*v;
            u.y = u.y + p0.y*v;
            X[k*nm + t] = u;
        }
    }
}

// Spectrum position m of a transformed axis from the modes: the mode index
// (or -1 for a zero) and the factor (inverse of fwdMode, with the
// reflection symmetries X_{2N-m} = -X_m (DCT-II), S_{2N-m} = S_m (DST-II)).
// Quarter-wave axes (L = 4N): x_a = 1/2N * 2 sum_k X_k cos or sin(phi_k
// (2a+1)) is the real inverse FFT, scaled by 1/4N, of the spectrum with
// 2 e^{i phi_k} X_k (DCT-IV; -2i e^{i phi_k} X_k for DST-IV) at 2k+1 and
// its conjugate factor (2 e^{-i phi_k}; +2i e^{-i phi_k}) at 4N-1-2k.
__device__ __forceinline__ void bwdMode(int code, int n, int m, int& k, C2& w)
{
    if (code == C_PER) { k = m; w = C2{1.0, 0.0}; return; }
    const double pi = 3.14159265358979323846;
    if (code == C_ND || code == C_DN)
    {
        if ((m & 1) == 0) { k = -1; return; }
        const bool lower = m < 2*n;
        k = lower ? (m - 1)/2 : (4*n - 1 - m)/2;
        const double phi = pi*(2*k + 1)/(4.0*n);
        const C2 e = cexpi(lower ? phi : -phi);
        if (code == C_ND)
            w = C2{2.0*e.x, 2.0*e.y};
        else
            w = lower ? C2{2.0*e.y, -2.0*e.x} : C2{-2.0*e.y, 2.0*e.x};
        return;
    }
    const C2 e = cexpi(pi*m/(2.0*n));
    if (code == C_NN)
    {
        if (m == n) { k = -1; return; }
        k = m < n ? m : 2*n - m;
        w = m < n ? e : C2{-e.x, -e.y};
        return;
    }
    if (m == 0) { k = -1; return; }
    k = m <= n ? m - 1 : 2*n - m - 1;
    w = C2{e.y, -e.x};   // -i*e
}

// Step 4: every position of the extended spectrum.
__global__ void k_fft_expand(FftGeom g, const C2* X, C2* E)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    const long plane = (long)g.L1*g.L2;
    if (t >= plane*g.ns) return;
    const int s = int(t/plane);
    const long r = t - (long)s*plane;
    const int m1 = int(r/g.L2), m2 = int(r % g.L2);
    int k1, k2;
    C2 w1, w2;
    bwdMode(g.code1, g.n1, m1, k1, w1);
    bwdMode(g.code2, g.n2, m2, k2, w2);
    C2 v{0.0, 0.0};
    if (k1 >= 0 && k2 >= 0)
        v = cmul(cmul(w1, w2),
                 X[((long)s*g.n1 + k1)*g.n2 + k2]);
    E[t] = v;
}

// Step 5: scaled real part into the interior of x.
__global__ void k_fft_out(FftGeom g, const C2* E, double* x, double scale)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)g.N[0]*g.N[1]*g.N[2]) return;
    int q[3];
    q[0] = int(t/((long)g.N[1]*g.N[2]));
    q[1] = int((t/g.N[2]) % g.N[1]);
    q[2] = int(t % g.N[2]);
    const long e = ((long)q[g.s]*g.L1 + q[g.d1])*g.L2 + q[g.d2];
    x[refIndex(g, q[0], q[1], q[2])] = E[e].x*scale;
}

// Every ghost of x (meshLevel::correctBoundaryConditions on level 0): face
// ghosts by their condition -- periodic image, Neumann x_P + g/delta with
// delta = 1/h of the mirrored boundary cell, Dirichlet 2v - x_P -- and edge
// and corner ghosts by periodic image when every face they border is
// periodic, else a copy of the adjacent interior cell.
__global__ void k_fft_ghosts(FftGeom g, double* x)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)g.P[0]*g.P[1]*g.P[2]) return;
    int q[3];
    q[0] = int(t/((long)g.P[1]*g.P[2]));
    q[1] = int((t/g.P[2]) % g.P[1]);
    q[2] = int(t % g.P[2]);
    int bo[3], nout = 0;
    bool allPer = true;
    for (int d = 0; d < 3; d++)
    {
        bo[d] = q[d] == 0 ? -1 : (q[d] == g.P[d] - 1 ? 1 : 0);
        nout += bo[d] != 0;
        if (bo[d] != 0 && g.bc[2*d + (bo[d] > 0)] != PER) allPer = false;
    }
    if (nout == 0) return;
    int sidx[3];
    for (int d = 0; d < 3; d++)
        sidx[d] = allPer ? q[d] - bo[d]*g.N[d] : q[d] - bo[d];
    const double xi = x[((long)sidx[0]*g.P[1] + sidx[1])*g.P[2] + sidx[2]];
    double v = xi;
    if (nout == 1 && !allPer)
    {
        const int d = bo[0] ? 0 : (bo[1] ? 1 : 2);
        const int f = 2*d + (bo[d] > 0);
        if (g.bc[f] == NEU)
        {
            const double delta = 1.0/g.hwall[f];
            v = xi + g.bcval[f]/delta;
        }
        else
            v = 2.0*g.bcval[f] - xi;
    }
    x[t] = v;
}

void rocfftCheck(rocfft_status s, const char* what)
{
    if (s != rocfft_status_success)
        throw Error(BMG_ERR_HIP, std::string("rocFFT ") + what + " failed (" +
                                     std::to_string(int(s)) + ")");
}

std::once_flag g_rocfftSetup;

}  // namespace

}  // namespace bmg

using namespace bmg;

struct bmg_fft
{
    int device = 0;
    hipStream_t stream = nullptr;
    FftGeom g{};
    C2* E = nullptr;       // extended spectrum, ns x L1 x L2
    C2* X = nullptr;       // modes, ns x n1 x n2
    double* Dh = nullptr;
    double* V = nullptr;
    double* h = nullptr;   // widths along s, and their east/west means
    double* de = nullptr;
    double* dw = nullptr;
    void* work = nullptr;
    rocfft_plan fwd = nullptr, bwd = nullptr;
    rocfft_execution_info info = nullptr;

    ~bmg_fft()
    {
        if (fwd) rocfft_plan_destroy(fwd);
        if (bwd) rocfft_plan_destroy(bwd);
        if (info) rocfft_execution_info_destroy(info);
        for (void* p : {(void*)E, (void*)X, (void*)Dh, (void*)V, (void*)h,
                        (void*)de, (void*)dw, work})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace bmg {

void fft_create(bmg_fft** out, const bmg_grid* grid)
{
    if (!out || !grid) throw Error(BMG_ERR_ARG, "null pointer");
    std::unique_ptr<bmg_fft> F(new bmg_fft);
    FftGeom& g = F->g;
    std::vector<double> w[3];
    bool uniform[3];
    for (int d = 0; d < 3; d++)
    {
        g.N[d] = grid->n[d];
        g.P[d] = g.N[d] + 2;
        if (g.N[d] < 1 || !grid->vertices[d])
            throw Error(BMG_ERR_ARG, "bad grid");
        w[d].resize(g.N[d]);
        for (int i = 0; i < g.N[d]; i++)
            w[d][i] = grid->vertices[d][i + 1] - grid->vertices[d][i];
        uniform[d] = true;
        for (int i = 0; i < g.N[d]; i++)
            if (std::fabs(w[d][i] - w[d][0]) > 1e-12*std::fabs(w[d][0]))
                uniform[d] = false;
    }
    for (int f = 0; f < 6; f++)
    {
        g.bc[f] = grid->bc[f];
        g.bcval[f] = grid->bc_value[f];
        g.hwall[f] = (f & 1) ? w[f/2].back() : w[f/2].front();
    }
    for (int f = 0; f < 6; f++)
        if ((g.bc[f] == PER) != (g.bc[f ^ 1] == PER))
            throw Error(BMG_ERR_ARG, "periodic faces come in pairs");
    // planner.C:76-117 (one process)
    int nonUniform = 0;
    g.s = -1;
    for (int d = 0; d < 3; d++)
        if (!uniform[d])
        {
            nonUniform++;
            if (g.s < 0) g.s = d;
        }
    if (nonUniform > 1)
        throw Error(BMG_ERR_ARG, "at least two mesh directions must be "
                                 "uniform for the FFT solver");
    if (g.s < 0) g.s = g.N[0] > 1 ? 0 : 2;
    g.d1 = g.s == 0 ? 2 : 0;
    g.d2 = g.s == 2 ? 1 : (g.s == 0 ? 1 : 2);
    g.n1 = g.N[g.d1];
    g.n2 = g.N[g.d2];
    g.ns = g.N[g.s];
    g.code1 = axisCode(g.bc[2*g.d1], g.bc[2*g.d1 + 1]);
    g.code2 = axisCode(g.bc[2*g.d2], g.bc[2*g.d2 + 1]);
    g.codes = axisCode(g.bc[2*g.s], g.bc[2*g.s + 1]);
    if (g.ns < 2 || (g.codes == C_PER && g.ns < 3))
        throw Error(BMG_ERR_ARG, "too few cells along the tridiagonal axis");
    auto extLen = [](int code, int n) {
        return code == C_PER ? n : (code == C_ND || code == C_DN ? 4*n : 2*n);
    };
    g.L1 = extLen(g.code1, g.n1);
    g.L2 = extLen(g.code2, g.n2);
    g.d1sqr = w[g.d1][0]*w[g.d1][0];
    g.d2sqr = w[g.d2][0]*w[g.d2][0];
    // widths along s with the ghost widths (mirrored, or the periodic
    // images) and their means (tridiagonalSolver.C:112-130)
    const std::vector<double>& ws = w[g.s];
    const int ns = g.ns;
    const bool per = g.codes == C_PER;
    std::vector<double> hg(ns + 2), h(ns), de(ns), dw(ns);
    hg[0] = per ? ws[ns - 1] : ws[0];
    hg[ns + 1] = per ? ws[0] : ws[ns - 1];
    for (int k = 0; k < ns; k++) hg[k + 1] = ws[k];
    for (int k = 0; k < ns; k++)
    {
        h[k] = hg[k + 1];
        de[k] = (hg[k + 1] + hg[k + 2])/2.0;
        dw[k] = (hg[k + 1] + hg[k])/2.0;
    }
    HIPCK(hipGetDevice(&F->device));
    HIPCK(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking));
    const size_t nE = size_t(ns)*g.L1*g.L2;
    const size_t nX = size_t(ns)*g.n1*g.n2;
    F->E = devAlloc<C2>(nE);
    F->X = devAlloc<C2>(nX);
    F->Dh = devAlloc<double>(nX);
    F->V = devAlloc<double>(per ? nX : 1);
    F->h = devUpload(h);
    F->de = devUpload(de);
    F->dw = devUpload(dw);
    std::call_once(g_rocfftSetup, [] { rocfftCheck(rocfft_setup(), "setup"); });
    const size_t lengths[2] = {size_t(g.L2), size_t(g.L1)};
    rocfftCheck(rocfft_plan_create(&F->fwd, rocfft_placement_inplace,
                                   rocfft_transform_type_complex_forward,
                                   rocfft_precision_double, 2, lengths,
                                   size_t(ns), nullptr), "plan");
    rocfftCheck(rocfft_plan_create(&F->bwd, rocfft_placement_inplace,
                                   rocfft_transform_type_complex_inverse,
                                   rocfft_precision_double, 2, lengths,
                                   size_t(ns), nullptr), "plan");
    size_t w1 = 0, w2 = 0;
    rocfftCheck(rocfft_plan_get_work_buffer_size(F->fwd, &w1), "work size");
    rocfftCheck(rocfft_plan_get_work_buffer_size(F->bwd, &w2), "work size");
    rocfftCheck(rocfft_execution_info_create(&F->info), "execution info");
    rocfftCheck(rocfft_execution_info_set_stream(F->info, F->stream),
                "stream");
    const size_t wb = std::max(w1, w2);
    if (wb)
    {
        HIPCK(hipMalloc(&F->work, wb));
        rocfftCheck(rocfft_execution_info_set_work_buffer(F->info, F->work,
                                                          wb), "work buffer");
    }
    *out = F.release();
}

void fft_destroy(bmg_fft* F) { delete F; }

// FFTPoissonSolver::solve (FFTPoissonSolver.C:97-465) on device arrays in
// the reference layout; b may be null (b = 0).  With dxold the time term:
// the source minus x_old/dT and every diagonal minus 1/dT, dT =
// deltaT*dtFrac (FFTPoissonSolver.C:116-131, tridiagonalSolver.C:334-352).
void fft_solve_device(bmg_fft* F, const double* db, double* dx,
                      const double* dxold, double dT)
{
    if (!F || !dx) throw Error(BMG_ERR_ARG, "null pointer");
    if (dxold && !(std::isfinite(dT) && dT > 0.0))
        throw Error(BMG_ERR_ARG, "time term needs deltaT*dtFrac > 0");
    FftGeom g = F->g;
    g.dT = dxold ? dT : 0.0;
    const long cells = long(g.N[0])*g.N[1]*g.N[2];
    const long padded = long(g.P[0])*g.P[1]*g.P[2];
    const long nE = long(g.ns)*g.L1*g.L2;
    const long nm = long(g.n1)*g.n2;
    hipStream_t s = F->stream;
    auto blocks = [](long n) { return dim3(unsigned((n + 255)/256)); };
    hipLaunchKernelGGL(k_fft_rhs, blocks(cells), dim3(256), 0, s, g, db, dx,
                       dxold, F->E);
    HIPCK(hipGetLastError());
    void* buf[1] = {F->E};
    rocfftCheck(rocfft_execute(F->fwd, buf, nullptr, F->info), "forward");
    hipLaunchKernelGGL(k_fft_tridiag, blocks(nm), dim3(256), 0, s, g, F->E,
                       F->X, F->Dh, F->V, F->h, F->de, F->dw);
    hipLaunchKernelGGL(k_fft_expand, blocks(nE), dim3(256), 0, s, g, F->X,
                       F->E);
    HIPCK(hipGetLastError());
    rocfftCheck(rocfft_execute(F->bwd, buf, nullptr, F->info), "inverse");
    const double scale = 1.0/(double(g.L1)*double(g.L2));
    hipLaunchKernelGGL(k_fft_out, blocks(cells), dim3(256), 0, s, g, F->E,
                       dx, scale);
    hipLaunchKernelGGL(k_fft_ghosts, blocks(padded), dim3(256), 0, s, g, dx);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(s));
}

void fft_solve(bmg_fft* F, const double* b, double* x, const double* xold,
               double dT)
{
    if (!F || !x) throw Error(BMG_ERR_ARG, "null pointer");
    const FftGeom& g = F->g;
    const size_t n = size_t(g.P[0])*g.P[1]*g.P[2];
    double* db = nullptr;
    double* dold = nullptr;
    double* dx = devAlloc<double>(n);
    try
    {
        if (b)
        {
            db = devAlloc<double>(n);
            HIPCK(hipMemcpy(db, b, n*sizeof(double), hipMemcpyHostToDevice));
        }
        if (xold)
        {
            dold = devAlloc<double>(n);
            HIPCK(hipMemcpy(dold, xold, n*sizeof(double),
                            hipMemcpyHostToDevice));
        }
        HIPCK(hipMemcpy(dx, x, n*sizeof(double), hipMemcpyHostToDevice));
        fft_solve_device(F, db, dx, dold, dT);
        HIPCK(hipMemcpy(x, dx, n*sizeof(double), hipMemcpyDeviceToHost));
    }
    catch (...)
    {
        devFree(db);
        devFree(dold);
        devFree(dx);
        throw;
    }
    devFree(db);
    devFree(dold);
    devFree(dx);
}

}  // namespace bmg
