// hipMG.H -- C++ face of the MI355X multigrid solver with the interface of
// Briscola's MG linear solver, for host code that calls it the way Briscola
// calls its solvers:
//
//   solver<stencil,scalar,colocated>::solve(linearSystem&, bool constMatrix)
//   (src/briscolaFiniteVolume/solvers/solver/solver.H:280-284,
//    solvers/MG/MG.H:135-139, MG.C:402-483)
//
// Header-only; everything numerical happens behind the C ABI of bmg.h.
//
// Mirrored semantics:
//   * the dictionary keys and defaults of MG (MG.C:257-304) and solver
//     (relTol and tolerance required, minIter 0, maxIter 999; solver.C:
//     114-123);
//   * solve() eliminates the Neumann/Dirichlet ghosts into A and the level-0
//     inhomogeneous terms into b, pins the coarsest level of singular systems
//     and overwrites x (MG.C:431-481); the mutated A and b are copied back
//     into the linearSystem, as the reference mutates the caller's objects;
//   * the coarse factorisation is redone unless constMatrix (MG.C:449-459);
//   * diagonal systems are solved by one diagonal sweep (MG.C:408-428);
//   * printSolverStats line "MG: Solving for colocated <name>, initial
//     residual = ..., final residual = ..., nIter = N" (solver.C:97-112);
//   * non-convergence is silent; configuration errors raise FatalError (the
//     reference aborts through FatalErrorInFunction).
#ifndef BRISCOLA_AMD_HIPMG_H
#define BRISCOLA_AMD_HIPMG_H

#include "bmg.h"

#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace briscola_amd
{

struct FatalError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

// The subset of an OpenFOAM dictionary MG reads: word/scalar/label entries.
class dictionary
{
    std::map<std::string, std::string> e_;

public:
    dictionary() = default;
    dictionary(std::initializer_list<std::pair<const std::string, std::string>> l)
    : e_(l) {}

    void add(const std::string& k, const std::string& v) { e_[k] = v; }
    bool found(const std::string& k) const { return e_.count(k) > 0; }

    double lookupScalar(const std::string& k) const
    {
        auto it = e_.find(k);
        if (it == e_.end())
            throw FatalError("keyword " + k + " is undefined in dictionary");
        return std::stod(it->second);
    }
    int lookupOrDefault(const std::string& k, int d) const
    {
        auto it = e_.find(k);
        return it == e_.end() ? d : std::stoi(it->second);
    }
    std::string lookupOrDefault(const std::string& k, const char* d) const
    {
        auto it = e_.find(k);
        return it == e_.end() ? std::string(d) : it->second;
    }
};

// What MG reads from fvMesh: level-0 shape, rectilinear vertex coordinates
// and the boundary condition of every face of x (-x,+x,-y,+y,-z,+z); the
// Poisson solver also reads the cell volumes (level 0, ghost-padded, as the
// mesh metrics hold them, fvMeshMetrics.C:62-100).
struct fvMesh
{
    int N[3];
    std::vector<double> vertices[3];
    int bc[6];                 // BMG_NEUMANN / BMG_DIRICHLET / BMG_PERIODIC
    double bcValue[6] = {0, 0, 0, 0, 0, 0};
    // per-cell data of a Neumann/Dirichlet face, replacing bcValue[f] when
    // not empty: boundaryValues()[0] / boundaryGradients()[0] of the face's
    // boundary condition of x (DirichletBoundaryConditionBase.H:82,
    // NeumannBoundaryConditionBase.H:82), the face's layer of level-0
    // cells with the normal axis of size 1, last axis fastest
    std::vector<double> bcValues[6];
    std::vector<double> cellVolumes;
    double deltaT = 0.0;       // time().deltaTValue() (ddtScheme.H:59)
    double deltaT0 = 0.0;      // time().deltaT0Value(), the previous step
};

// linearSystem<stencil,scalar,colocated>: unknown x, deep stencil field A
// (every level, ghost-padded [C,L,R,B,T,A,F]) and source b (level 0).
struct linearSystem
{
    std::string name = "p";
    std::vector<std::vector<double>> A;   // A[level]: (Nx+2)(Ny+2)(Nz+2)*7
    std::vector<double> b;                // level 0, ghost-padded
    std::vector<double> x;                // level 0, ghost-padded
    bool symmetric = true;
    bool diagonal = false;
};

struct solverStats
{
    int nIter = 0;
    double initialResidual = 0.0;
    double finalResidual = 0.0;
};

class solver
{
public:
    virtual ~solver() = default;
    virtual void solve(linearSystem& sys, bool constMatrix = false) = 0;
};

inline void check(int code)
{
    if (code != BMG_OK) throw FatalError(bmg_last_error());
}

// Solver options from the MG dictionary keys and defaults (MG.C:257-304,
// solver.C:114-123), plus this solver's "precision" (FP64 | mixed),
// "krylov" (none | FCG: flexible CG preconditioned by one cycle) and
// "deflation" (off | on: FCG deflated by the high-coefficient inclusions).
inline bmg_opts optsFromDict(const dictionary& dict)
{
    bmg_opts o;
    bmg_default_opts(&o);
    o.rel_tol = dict.lookupScalar("relTol");
    o.tolerance = dict.lookupScalar("tolerance");
    o.min_iter = dict.lookupOrDefault("minIter", 0);
    o.max_iter = dict.lookupOrDefault("maxIter", 999);
    o.n_sweeps_pre = dict.lookupOrDefault("nSweepsPre", 0);
    o.n_sweeps_post = dict.lookupOrDefault("nSweepsPost", 2);
    const std::string cyc = dict.lookupOrDefault("cycleType", "F");
    if (cyc == "V") o.cycle = BMG_VCYCLE;
    else if (cyc == "W") o.cycle = BMG_WCYCLE;
    else if (cyc == "F") o.cycle = BMG_FCYCLE;
    else throw FatalError("Unknown cycleType " + cyc);
    const std::string cm = dict.lookupOrDefault("coarseMode", "direct");
    if (cm == "direct") o.coarse_mode = BMG_COARSE_DIRECT;
    else if (cm == "smooth") o.coarse_mode = BMG_COARSE_SMOOTH;
    else throw FatalError("Unknown coarseMode " + cm);
    o.coarse_level = dict.lookupOrDefault("coarseLevel", 0);
    const std::string sm = dict.lookupOrDefault("smoother", "rbgs");
    if (sm == "rbgs") o.smoother = BMG_RBGS;
    else if (sm == "jac") o.smoother = BMG_JAC;
    else if (sm == "lexgs") o.smoother = BMG_LEXGS;
    else throw FatalError("Unsupported smoother type " + sm);
    if (dict.lookupOrDefault("prolong", "linear") != "linear" ||
        dict.lookupOrDefault("restrict", "linear") != "linear")
        throw FatalError("only the linear transfer schemes are provided");
    const std::string prec = dict.lookupOrDefault("precision", "FP64");
    o.precision = prec == "mixed" ? BMG_MIXED : BMG_FP64;
    const std::string kry = dict.lookupOrDefault("krylov", "none");
    if (kry == "FCG") o.krylov = BMG_KRYLOV_FCG;
    else if (kry != "none") throw FatalError("Unknown krylov " + kry);
    const std::string defl = dict.lookupOrDefault("deflation", "off");
    if (defl != "on" && defl != "off")
        throw FatalError("deflation must be on or off");
    o.deflation = defl == "on" ? 1 : 0;
    return o;
}

// A solver context on the mesh's level 0 (vertices copied into verts).
inline bmg_ctx* createContext(const fvMesh& mesh, const bmg_opts& opts,
                              std::vector<double> (&verts)[3])
{
    bmg_grid g;
    for (int d = 0; d < 3; d++)
    {
        g.n[d] = mesh.N[d];
        verts[d] = mesh.vertices[d];
        g.vertices[d] = verts[d].data();
    }
    for (int f = 0; f < 6; f++)
    {
        g.bc[f] = mesh.bc[f];
        g.bc_value[f] = mesh.bcValue[f];
    }
    bmg_ctx* ctx = nullptr;
    check(bmg_create(&ctx, &g, &opts));
    return ctx;
}

// The per-cell boundary data of the mesh's faces into the context (read
// again before every solve: boundary values may change between solves).
inline void applyBoundaryData(bmg_ctx* ctx, const fvMesh& mesh)
{
    for (int f = 0; f < 6; f++)
    {
        if (mesh.bc[f] == BMG_PERIODIC) continue;
        const int d = f/2;
        const size_t n = size_t(mesh.N[d == 0 ? 1 : 0])*mesh.N[d == 2 ? 1 : 2];
        if (mesh.bcValues[f].empty())
            check(bmg_set_boundary_values(ctx, f, nullptr));
        else if (mesh.bcValues[f].size() != n)
            throw FatalError("boundary data of face " + std::to_string(f) +
                             " must hold one value per boundary cell");
        else
            check(bmg_set_boundary_values(ctx, f, mesh.bcValues[f].data()));
    }
}

class MG : public solver
{
    bmg_ctx* ctx_ = nullptr;
    bmg_opts opts_;
    std::vector<double> verts_[3];
    bool operatorSet_ = false;
    bool verbose_;
    solverStats last_;
    const fvMesh* mesh_;
    int bc_[6];

public:
    static constexpr const char* typeName = "MG";

    // MG(const dictionary&, const fvMesh&) -- MG.C:257-399
    MG(const dictionary& dict, const fvMesh& mesh, bool verbose = true)
    : opts_(optsFromDict(dict)), verbose_(verbose), mesh_(&mesh)
    {
        ctx_ = createContext(mesh, opts_, verts_);
        for (int f = 0; f < 6; f++) bc_[f] = mesh.bc[f];
    }

    ~MG() override { if (ctx_) bmg_destroy(ctx_); }
    MG(const MG&) = delete;
    MG& operator=(const MG&) = delete;

    int nLevels() const { return bmg_num_levels(ctx_); }

    std::vector<int> levelShape(int l) const
    {
        int n[3];
        check(bmg_level_shape(ctx_, l, n));
        return {n[0], n[1], n[2]};
    }

    const solverStats& lastStats() const { return last_; }

    // MG::solve (MG.C:402-483)
    void solve(linearSystem& sys, bool constMatrix = false) override
    {
        const int L = nLevels();
        applyBoundaryData(ctx_, *mesh_);
        if (sys.diagonal)
        {
            check(bmg_solve_diagonal(ctx_, sys.A.at(0).data(), sys.b.data(),
                                     sys.x.data()));
            last_ = {0, 1.0, 0.0};
            print("diagonal", sys.name);
            return;
        }
        if (int(sys.A.size()) != L)
            throw FatalError("linearSystem must hold A on every level");
        if (!constMatrix || !operatorSet_)
        {
            for (int l = 0; l < L; l++)
                check(bmg_set_operator_stencil(ctx_, l, sys.A[l].data(),
                                               sys.symmetric));
            operatorSet_ = true;
        }
        else if (hasGhostCoefficients(sys))
            // a newly assembled system with the same matrix: its ghost
            // coefficients are still in place, so MG::solve's eliminateGhosts
            // folds the boundary data into b again (MG.C:445)
            check(bmg_renew_source_terms(ctx_));
        bmg_stats st;
        check(bmg_solve(ctx_, sys.b.data(), sys.x.data(), constMatrix, &st));
        // the reference eliminates ghosts and pins the coarse level in place
        for (int l = 0; l < L; l++)
            check(bmg_get_operator_stencil(ctx_, l, sys.A[l].data()));
        last_ = {st.iters, st.initial_residual, st.final_residual};
        print(typeName, sys.name);
    }

private:
    // Whether level 0 of sys.A still couples a boundary cell to a ghost
    // across a Neumann or Dirichlet face (elimination zeroes these,
    // NeumannBoundaryCondition.C:48, DirichletBoundaryCondition.C:42).
    bool hasGhostCoefficients(const linearSystem& sys) const
    {
        const std::vector<int> n = levelShape(0);
        const long P1 = n[1] + 2, P2 = n[2] + 2;
        const std::vector<double>& A = sys.A.at(0);
        for (int f = 0; f < 6; f++)
        {
            if (bc_[f] == BMG_PERIODIC) continue;
            const int d = f/2;
            const int a1 = d == 0 ? 1 : 0, a2 = d == 2 ? 1 : 2;
            for (int t1 = 1; t1 <= n[a1]; t1++)
                for (int t2 = 1; t2 <= n[a2]; t2++)
                {
                    int c[3];
                    c[d] = (f & 1) ? n[d] : 1;
                    c[a1] = t1;
                    c[a2] = t2;
                    const long p = (c[0]*P1 + c[1])*P2 + c[2];
                    if (A[7*p + f + 1] != 0.0) return true;
                }
        }
        return false;
    }

    void print(const char* solverName, const std::string& field) const
    {
        if (!verbose_) return;
        std::printf("%s: Solving for colocated %s, initial residual = %g, "
                    "final residual = %g, nIter = %d\n", solverName,
                    field.c_str(), last_.initialResidual, last_.finalResidual,
                    last_.nIter);
    }
};

}  // namespace briscola_amd

#endif
