// hipSplitPoissonSolver.H -- C++ faces of Briscola's FFTPoissonSolver and
// splitPoissonSolver over the MI355X solvers, for host code that selects
// them where it would select the reference's:
//
//   FFTPoissonSolver<stencil>::solve(x, bPtr, lambdaPtr, ddt, dtFrac)
//   (src/briscolaFiniteVolume/solvers/PoissonSolvers/FFTPoissonSolver/
//    FFTPoissonSolver.C:85-458)
//   splitPoissonSolver<stencil,scalar,colocated>::solve(...)
//   (splitPoissonSolver/splitPoissonSolver.C:14-149)
//
// Header-only; solves, sources and fluxes run behind the C ABI of bmg.h.
#ifndef BRISCOLA_AMD_HIPSPLITPOISSONSOLVER_H
#define BRISCOLA_AMD_HIPSPLITPOISSONSOLVER_H

#include "hipPoissonSolver.H"

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

namespace briscola_amd
{

// FFTPoissonSolver: div(grad x) = -b in finite-difference form on a mesh
// with two uniform directions; no lambda (a FatalError, as in the
// reference); ddt adds the Euler time term with x.oldTime() = xOld; the flux
// is faceGrad(x)*faceAreas (FFTPoissonSolver.C:446-454).
class FFTPoissonSolver
{
    const fvMesh& mesh_;
    bmg_fft* fft_ = nullptr;
    bmg_ctx* geo_ = nullptr;        // level-0 metrics for the flux
    std::vector<double> verts_[3];
    bool computeFlux_ = false;
    faceField flux_;
    bool verbose_;

    size_t cells() const
    {
        return size_t(mesh_.N[0] + 2)*(mesh_.N[1] + 2)*(mesh_.N[2] + 2);
    }

public:
    FFTPoissonSolver(const dictionary&, const fvMesh& mesh, bool verbose = true)
    : mesh_(mesh), verbose_(verbose)
    {
        bmg_grid g;
        for (int d = 0; d < 3; d++)
        {
            g.n[d] = mesh.N[d];
            verts_[d] = mesh.vertices[d];
            g.vertices[d] = verts_[d].data();
        }
        for (int f = 0; f < 6; f++)
        {
            g.bc[f] = mesh.bc[f];
            g.bc_value[f] = mesh.bcValue[f];
        }
        check(bmg_fft_create(&fft_, &g));
    }

    ~FFTPoissonSolver()
    {
        if (fft_) bmg_fft_destroy(fft_);
        if (geo_) bmg_destroy(geo_);
    }
    FFTPoissonSolver(const FFTPoissonSolver&) = delete;
    FFTPoissonSolver& operator=(const FFTPoissonSolver&) = delete;

    void enableFlux()
    {
        computeFlux_ = true;
        if (!geo_)
        {
            bmg_opts o;
            bmg_default_opts(&o);
            std::vector<double> v[3];
            geo_ = createContext(mesh_, o, v);
        }
    }
    void disableFlux()
    {
        computeFlux_ = false;
        for (auto& v : flux_.f) v.clear();
    }
    bool computeFlux() const { return computeFlux_; }
    const faceField& flux() const
    {
        if (flux_.f[0].empty()) throw FatalError("Flux not set");
        return flux_;
    }

    void solve(meshField& x, const meshField* bPtr,
               const faceField* lambdaPtr, bool ddt = false,
               double dtFrac = 1.0, const RungeKuttaScheme* = nullptr,
               const meshField* xOld = nullptr)
    {
        const size_t n = cells();
        if (lambdaPtr)
            throw FatalError("FFT Poisson solver does not support use of "
                             "lambda argument in solve function.");
        if (x.size() != n) throw FatalError("x must be the padded level 0");
        if (bPtr && bPtr->size() != n) throw FatalError("b size");
        const double* b = bPtr ? bPtr->data() : nullptr;
        if (ddt)
        {
            if (!xOld || xOld->size() != n)
                throw FatalError("ddt needs the old-time field of x");
            check(bmg_fft_solve_ddt(fft_, b, x.data(), xOld->data(),
                                    mesh_.deltaT, dtFrac));
        }
        else
            check(bmg_fft_solve(fft_, b, x.data()));
        if (verbose_)
            std::printf("FFT: Solving for colocated p, residual = 0, "
                        "nIter = 1\n");
        if (computeFlux_)
        {
            for (auto& v : flux_.f) v.assign(n, 0.0);
            check(bmg_face_flux(geo_, x.data(), nullptr, nullptr, nullptr,
                                flux_.f[0].data(), flux_.f[1].data(),
                                flux_.f[2].data()));
        }
    }
};

// splitPoissonSolver: div(lambda grad x) = -b solved by constant-
// coefficient solves on the modified source bHat = b/lambda0 +
// div((lambda/lambda0 - 1) faceGrad(x) A), lambda0 the largest face
// coefficient, nCorr times, from the prediction (1 + r) x_old - r x_oldOld
// with r = deltaT*dtFrac/deltaT0 and after the optional preconditioner
// solve (splitPoissonSolver.C:53-149).  The inner solver is the FFT solver
// (type FFT) or the multigrid one (any other type); the preconditioner is
// the multigrid Poisson solver.
class splitPoissonSolver
{
    const fvMesh& mesh_;
    bmg_ctx* geo_ = nullptr;
    std::vector<double> verts_[3];
    std::unique_ptr<PoissonSolver> mg_;
    std::unique_ptr<FFTPoissonSolver> fft_;
    std::unique_ptr<PoissonSolver> pre_;
    int nCorr_;
    bool prediction_;
    bool computeFlux_ = false;
    faceField flux_;
    meshField bHat_;
    faceField phi_;

    size_t cells() const
    {
        return size_t(mesh_.N[0] + 2)*(mesh_.N[1] + 2)*(mesh_.N[2] + 2);
    }

    static bool readSwitch(const dictionary& d, const std::string& k, bool def)
    {
        const std::string v = d.lookupOrDefault(k, def ? "true" : "false");
        if (v == "true" || v == "on" || v == "yes" || v == "1") return true;
        if (v == "false" || v == "off" || v == "no" || v == "0") return false;
        throw FatalError("bad Switch value " + v + " for " + k);
    }

    // the largest lambda over the faces of every direction (faces 1..N+1
    // along the direction, interior cells along the others)
    double lambda0(const faceField& lam) const
    {
        const int P1 = mesh_.N[1] + 2, P2 = mesh_.N[2] + 2;
        double m = 0.0;
        for (int d = 0; d < 3; d++)
            for (int i = 1; i <= mesh_.N[0] + (d == 0); i++)
                for (int j = 1; j <= mesh_.N[1] + (d == 1); j++)
                    for (int k = 1; k <= mesh_.N[2] + (d == 2); k++)
                        m = std::max(m, lam.f[d][(size_t(i)*P1 + j)*P2 + k]);
        return m;
    }

public:
    splitPoissonSolver(const dictionary& dict, const dictionary& solverDict,
                       const fvMesh& mesh,
                       const dictionary* preconditionerDict = nullptr,
                       bool verbose = true)
    : mesh_(mesh), nCorr_(dict.lookupOrDefault("nCorr", 1)),
      prediction_(readSwitch(dict, "prediction", true))
    {
        if (nCorr_ < 1) throw FatalError("nCorr must be at least 1");
        if (solverDict.lookupOrDefault("type", "MG") == "FFT")
            fft_.reset(new FFTPoissonSolver(solverDict, mesh, verbose));
        else
            mg_.reset(new PoissonSolver(solverDict, mesh, verbose));
        if (preconditionerDict)
        {
            pre_.reset(new PoissonSolver(*preconditionerDict, mesh, verbose));
            pre_->enableFlux();
        }
        bmg_opts o;
        bmg_default_opts(&o);
        geo_ = createContext(mesh, o, verts_);
    }

    ~splitPoissonSolver() { if (geo_) bmg_destroy(geo_); }
    splitPoissonSolver(const splitPoissonSolver&) = delete;
    splitPoissonSolver& operator=(const splitPoissonSolver&) = delete;

    void enableFlux()
    {
        computeFlux_ = true;
        fft_ ? fft_->enableFlux() : mg_->enableFlux();
    }
    bool computeFlux() const { return computeFlux_; }
    const faceField& flux() const
    {
        if (flux_.f[0].empty()) throw FatalError("Flux not set");
        return flux_;
    }
    const meshField& bHat() const { return bHat_; }

    // xOld, xOldOld: x.oldTime() and x.oldTime().oldTime() (the prediction
    // reads both; ddt reads xOld).  rk: the flow solver's Runge-Kutta scheme
    // (the reference finds it in the object registry as "rkScheme"); the
    // preconditioner and a multigrid inner solver are defaultPoissonSolvers,
    // which then start from, and store, their own stage solution
    // (defaultPoissonSolver.C:67-80,97-104) -- so the prediction and the
    // preconditioner's result are replaced by the inner solver's stored
    // stage before its first solve, as in the reference.  The FFT inner
    // solver has no stages.
    void solve(meshField& x, const meshField* bPtr,
               const faceField* lambdaPtr, bool ddt = false,
               double dtFrac = 1.0, const meshField* xOld = nullptr,
               const meshField* xOldOld = nullptr,
               const RungeKuttaScheme* rk = nullptr)
    {
        const size_t n = cells();
        if (!lambdaPtr)
            throw FatalError("Poisson equation has no coefficient so it "
                             "cannot be split.");
        if (x.size() != n) throw FatalError("x must be the padded level 0");
        for (const auto& a : lambdaPtr->f)
            if (a.size() != n) throw FatalError("lambda size");
        if (bPtr && bPtr->size() != n) throw FatalError("b size");

        // prediction (splitPoissonSolver.C:69-80)
        if (prediction_)
        {
            if (!xOld || !xOldOld || xOld->size() != n || xOldOld->size() != n)
                throw FatalError("the prediction needs x.oldTime() and "
                                 "x.oldTime().oldTime()");
            if (!(mesh_.deltaT0 > 0.0))
                throw FatalError("the prediction needs deltaT0 > 0");
            const double deltaT = mesh_.deltaT*dtFrac;
            const double r = deltaT/mesh_.deltaT0;
            for (size_t c = 0; c < n; c++)
            {
                const double a = (1.0 + r)*(*xOld)[c];
                const double b = r*(*xOldOld)[c];
                x[c] = a - b;
            }
        }

        // preconditioner (splitPoissonSolver.C:84-87)
        if (pre_) pre_->solve(x, bPtr, lambdaPtr, ddt, dtFrac, rk, xOld);

        const double l0 = lambda0(*lambdaPtr);
        bHat_.assign(n, 0.0);
        for (auto& v : phi_.f) v.assign(n, 0.0);
        for (int corr = 0; corr < nCorr_; corr++)
        {
            check(bmg_split_source(geo_, x.data(), lambdaPtr->f[0].data(),
                                   lambdaPtr->f[1].data(),
                                   lambdaPtr->f[2].data(), l0,
                                   bPtr ? bPtr->data() : nullptr,
                                   bHat_.data(), phi_.f[0].data(),
                                   phi_.f[1].data(), phi_.f[2].data()));
            if (fft_) fft_->solve(x, &bHat_, nullptr, ddt, dtFrac, nullptr, xOld);
            else mg_->solve(x, &bHat_, nullptr, ddt, dtFrac, rk, xOld);
            // flux = lambda0*(inner flux + phi) (splitPoissonSolver.C:
            // 141-142)
            if (computeFlux_ && corr == nCorr_ - 1)
            {
                const faceField& f = fft_ ? fft_->flux() : mg_->flux();
                for (int d = 0; d < 3; d++)
                {
                    flux_.f[d].assign(n, 0.0);
                    for (size_t c = 0; c < n; c++)
                        flux_.f[d][c] = l0*(f.f[d][c] + phi_.f[d][c]);
                }
            }
        }
    }
};

}  // namespace briscola_amd

#endif
