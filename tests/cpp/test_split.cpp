// Drives the split Poisson solver adapter (briscola_amd/host/
// hipSplitPoissonSolver.H) the way a two-phase flow solver drives
// splitPoissonSolver: lambda faces, a source, an inner FFT or multigrid
// solver, nCorr corrections, optional prediction and time term, the flux.
//
//   test_split <dir>
// reads  dir/shape.txt  "nx ny nz bc0..bc5 inner nCorr prediction ddt
//                        deltaT deltaT0 dtFrac relTol precond
//                        [bcValue0..5]"
//        (inner: FFT or MG; precond 1: a multigrid preconditioner solve), dir/vertices.bin, dir/volumes.bin,
//        dir/lambda.bin (3 face arrays), dir/src.bin, dir/x0.bin,
//        dir/xold.bin, dir/xoldold.bin (padded level-0 fields)
// writes dir/x.bin, dir/flux.bin (3 arrays), dir/bhat.bin
#include "hipSplitPoissonSolver.H"

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

using namespace briscola_amd;

static std::vector<double> readBin(const std::string& p, size_t n)
{
    std::vector<double> v(n);
    std::ifstream f(p, std::ios::binary);
    f.read(reinterpret_cast<char*>(v.data()), n*sizeof(double));
    if (!f) throw FatalError("cannot read " + p);
    return v;
}

static void writeBin(const std::string& p, const std::vector<double>& v)
{
    std::ofstream f(p, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), v.size()*sizeof(double));
}

int main(int argc, char** argv)
{
    if (argc != 2) { std::fprintf(stderr, "usage: test_split <dir>\n"); return 2; }
    const std::string dir = argv[1];
    std::ifstream sh(dir + "/shape.txt");
    fvMesh mesh;
    std::string inner, prediction, relTol;
    int nCorr = 1, ddt = 0, precond = 0;
    double dtFrac = 1.0;
    sh >> mesh.N[0] >> mesh.N[1] >> mesh.N[2];
    for (int f = 0; f < 6; f++) sh >> mesh.bc[f];
    sh >> inner >> nCorr >> prediction >> ddt >> mesh.deltaT >> mesh.deltaT0
       >> dtFrac >> relTol >> precond;
    if (!sh) { std::fprintf(stderr, "bad shape.txt\n"); return 2; }
    // optional: the six uniform boundary values (Neumann gradient /
    // Dirichlet value per face)
    double bv[6];
    if (sh >> bv[0] >> bv[1] >> bv[2] >> bv[3] >> bv[4] >> bv[5])
        for (int f = 0; f < 6; f++) mesh.bcValue[f] = bv[f];
    const size_t nv = size_t(mesh.N[0] + 1) + mesh.N[1] + 1 + mesh.N[2] + 1;
    std::vector<double> v = readBin(dir + "/vertices.bin", nv);
    size_t o = 0;
    for (int d = 0; d < 3; d++)
    {
        mesh.vertices[d].assign(v.begin() + o, v.begin() + o + mesh.N[d] + 1);
        o += mesh.N[d] + 1;
    }
    const size_t n = size_t(mesh.N[0] + 2)*(mesh.N[1] + 2)*(mesh.N[2] + 2);
    mesh.cellVolumes = readBin(dir + "/volumes.bin", n);
    faceField lambda;
    std::vector<double> l = readBin(dir + "/lambda.bin", 3*n);
    for (int d = 0; d < 3; d++)
        lambda.f[d].assign(l.begin() + d*n, l.begin() + (d + 1)*n);
    const meshField b = readBin(dir + "/src.bin", n);
    meshField x = readBin(dir + "/x0.bin", n);
    const meshField xOld = readBin(dir + "/xold.bin", n);
    const meshField xOldOld = readBin(dir + "/xoldold.bin", n);
    dictionary dict{{"nCorr", std::to_string(nCorr)},
                    {"prediction", prediction}};
    dictionary solverDict{{"type", inner}, {"relTol", relTol},
                          {"tolerance", "0"}, {"cycleType", "V"},
                          {"nSweepsPre", "2"}, {"nSweepsPost", "2"},
                          {"coarseLevel", "1"}};
    try
    {
        dictionary preDict{{"type", "MG"}, {"relTol", relTol},
                           {"tolerance", "0"}, {"cycleType", "V"},
                           {"nSweepsPre", "2"}, {"nSweepsPost", "2"},
                           {"coarseLevel", "1"}};
        splitPoissonSolver ps(dict, solverDict, mesh,
                              precond ? &preDict : nullptr);
        ps.enableFlux();
        ps.solve(x, &b, &lambda, ddt != 0, dtFrac, &xOld, &xOldOld);
        writeBin(dir + "/x.bin", x);
        writeBin(dir + "/bhat.bin", ps.bHat());
        std::vector<double> f;
        for (int d = 0; d < 3; d++)
            f.insert(f.end(), ps.flux().f[d].begin(), ps.flux().f[d].end());
        writeBin(dir + "/flux.bin", f);
        // no coefficient: nothing to split
        bool threw = false;
        try { ps.solve(x, &b, nullptr); }
        catch (const FatalError&) { threw = true; }
        if (!threw) throw FatalError("split without lambda was accepted");
    }
    catch (const FatalError& e)
    {
        std::fprintf(stderr, "FatalError: %s\n", e.what());
        return 1;
    }
    return 0;
}
