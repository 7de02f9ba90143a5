// hipPoissonSolver.H -- C++ face of Briscola's defaultPoissonSolver over
// the MI355X multigrid, for host code that solves its pressure equation the
// way Briscola does:
//
//   PoissonSolver<stencil,scalar,colocated>::solve(x, bPtr, lambdaPtr,
//                                                  ddt, dtFrac)
//   (src/briscolaFiniteVolume/solvers/PoissonSolvers/PoissonSolver/
//    PoissonSolver.H:120-133, defaultPoissonSolver/defaultPoissonSolver.C:
//    39-124)
//
// Header-only; assembly, solve and flux run behind the C ABI of bmg.h.
//
// Mirrored semantics of defaultPoissonSolver::solve:
//   * the system is laplacian(lambda, x) (assembled on the device on every
//     level, lambda restricted by face area) plus the source: sys += b means
//     b_sys = 0 - V*b per cell (linearSystem.C:546-556);
//   * with ddt the Euler time term is subtracted, sys -= ddt(x)/dtFrac
//     (defaultPoissonSolver.C:91-92, EulerDdtScheme.C:50-54): the matrix
//     centre by (V/deltaT)/dtFrac on every level (assembled on the device)
//     and the source by ((V*xOld)/deltaT)/dtFrac, xOld being x.oldTime()
//     and deltaT the mesh's time step;
//   * constMatrix = !ddt && !lambda: without lambda or ddt the operator is
//     assembled once and the coarse factorisation kept (MG.C:449-459);
//   * with a Runge-Kutta scheme the previous solution of the same stage is
//     the initial guess, and the new solution is stored for the next time
//     step (the stages start as copies of x, defaultPoissonSolver.C:73-107);
//   * with the flux switched on, flux = lambda*faceGrad(x)*faceAreas on
//     every face after the solve (defaultPoissonSolver.C:108-121);
//   * the solver is the one the dictionary names (its keys are MG's; see
//     hipMG.H), and its statistics line is printed as MG prints it.
#ifndef BRISCOLA_AMD_HIPPOISSONSOLVER_H
#define BRISCOLA_AMD_HIPPOISSONSOLVER_H

#include "hipMG.H"

#include <string>
#include <vector>

namespace briscola_amd
{

// The Runge-Kutta scheme PoissonSolver looks up as "rkScheme": number of
// stages and the current stage (1-based).
struct RungeKuttaScheme
{
    int nStages = 1;
    int stage = 1;
};

// Level-0 cell field, ghost-padded (Nx+2)(Ny+2)(Nz+2), z contiguous.
using meshField = std::vector<double>;

// Face field of a colocated mesh: per axis d the lower face of every padded
// cell (the layout bmg_set_operator_laplacian takes for lambda).
struct faceField
{
    std::vector<double> f[3];
};

class PoissonSolver
{
    const fvMesh& mesh_;
    bmg_ctx* ctx_ = nullptr;
    bmg_opts opts_;
    std::vector<double> verts_[3];
    bool assembled_ = false;       // lambda-free operator assembled
    bool computeFlux_ = false;
    faceField flux_;
    std::vector<meshField> stages_;
    bool verbose_;
    solverStats last_;

    size_t cells() const
    {
        return size_t(mesh_.N[0] + 2)*(mesh_.N[1] + 2)*(mesh_.N[2] + 2);
    }

public:
    // PoissonSolver(const dictionary&, const fvMesh&)
    PoissonSolver(const dictionary& dict, const fvMesh& mesh,
                  bool verbose = true)
    : mesh_(mesh), opts_(optsFromDict(dict)), verbose_(verbose)
    {
        ctx_ = createContext(mesh, opts_, verts_);
    }

    ~PoissonSolver() { if (ctx_) bmg_destroy(ctx_); }
    PoissonSolver(const PoissonSolver&) = delete;
    PoissonSolver& operator=(const PoissonSolver&) = delete;

    void enableFlux() { computeFlux_ = true; }
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
    const solverStats& lastStats() const { return last_; }
    const std::vector<meshField>& stages() const { return stages_; }

    // defaultPoissonSolver::solve (defaultPoissonSolver.C:39-124).  xOld is
    // x.oldTime(), read when ddt is set.
    void solve(meshField& x, const meshField* bPtr,
               const faceField* lambdaPtr, bool ddt = false,
               double dtFrac = 1.0, const RungeKuttaScheme* rk = nullptr,
               const meshField* xOld = nullptr)
    {
        const size_t n = cells();
        if (x.size() != n) throw FatalError("x must be the padded level 0");
        if (ddt)
        {
            if (!xOld || xOld->size() != n)
                throw FatalError("ddt needs the old-time field of x, padded "
                                 "level 0");
            if (mesh_.cellVolumes.size() != n)
                throw FatalError("ddt needs the padded level-0 cell volumes");
            if (!(mesh_.deltaT > 0.0))
                throw FatalError("ddt needs deltaT > 0");
        }
        if (rk && stages_.empty()) stages_.assign(rk->nStages, x);

        // sys = im::laplacian(lambda, x), minus im::ddt(x)/dtFrac's matrix
        const double* l[3] = {nullptr, nullptr, nullptr};
        if (lambdaPtr)
            for (int d = 0; d < 3; d++)
            {
                if (lambdaPtr->f[d].size() != n)
                    throw FatalError("lambda size");
                l[d] = lambdaPtr->f[d].data();
            }
        if (ddt)
            check(bmg_set_operator_laplacian_ddt(ctx_, l[0], l[1], l[2],
                                                 mesh_.deltaT, dtFrac));
        else if (lambdaPtr || !assembled_)
            check(bmg_set_operator_laplacian(ctx_, l[0], l[1], l[2]));
        // a lambda or time-term operator replaces the plain one
        assembled_ = !lambdaPtr && !ddt;

        // sys += b: b_sys = 0 - V*b (linearSystem.C:546-556)
        std::vector<double> bsys(n, 0.0);
        if (bPtr)
        {
            if (bPtr->size() != n || mesh_.cellVolumes.size() != n)
                throw FatalError("b and the cell volumes must be padded "
                                 "level-0 fields");
            for (size_t c = 0; c < n; c++)
            {
                const double t = mesh_.cellVolumes[c]*(*bPtr)[c];
                bsys[c] = 0.0 - t;
            }
        }
        // sys -= ddt(x)/dtFrac: b_sys -= ((V*xOld)/deltaT)/dtFrac
        // (EulerDdtScheme.C:54, linearSystemI.H:392-410, linearSystem.C:
        // 694-706)
        if (ddt)
            for (size_t c = 0; c < n; c++)
            {
                const double t = mesh_.cellVolumes[c]*(*xOld)[c];
                bsys[c] = bsys[c] - (t/mesh_.deltaT)/dtFrac;
            }
        const bool constMatrix = !ddt && !lambdaPtr;

        if (rk) x = stages_.at(rk->stage - 1);

        // the system is assembled anew on every call (sys =
        // im::laplacian(...), defaultPoissonSolver.C:89), so MG::solve folds
        // the boundary data into b every time, also when the matrix is kept
        applyBoundaryData(ctx_, mesh_);
        check(bmg_renew_source_terms(ctx_));
        bmg_stats st;
        check(bmg_solve(ctx_, bsys.data(), x.data(), constMatrix, &st));
        last_ = {st.iters, st.initial_residual, st.final_residual};
        if (verbose_)
            std::printf("MG: Solving for colocated p, initial residual = %g, "
                        "final residual = %g, nIter = %d\n",
                        last_.initialResidual, last_.finalResidual,
                        last_.nIter);

        if (rk) stages_.at(rk->stage - 1) = x;

        if (computeFlux_)
        {
            for (auto& v : flux_.f) v.assign(n, 0.0);
            const double* l[3] = {nullptr, nullptr, nullptr};
            if (lambdaPtr)
                for (int d = 0; d < 3; d++) l[d] = lambdaPtr->f[d].data();
            check(bmg_face_flux(ctx_, x.data(), l[0], l[1], l[2],
                                flux_.f[0].data(), flux_.f[1].data(),
                                flux_.f[2].data()));
        }
    }
};

}  // namespace briscola_amd

#endif
