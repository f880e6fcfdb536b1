// poisson2_example -- the call sequence of G+Smo's examples/poisson2_example.cpp
// (gsExprAssembler + CG with diagonal preconditioning + L2 error per
// refinement level) on the drop-in classes of gismo.h.  The domain is
// built with gsNurbsCreator instead of read from an XML file:
//   -d 2|3        unit square / unit cube (default 2)
//   --grid NX NY  NX x NY unit squares glued at their interfaces (2D)
//   -f FILE       G+Smo XML file holding the (multi)patch geometry
//   -e N          degree elevation steps (default 1: degree 2)
//   -r N          refinement loops (default 5)
//   --last        solve only on the finest level
//   --full        gsPoissonAssembler (full pattern) instead of gsExprAssembler
//   --mixed       FP32 iterations with FP64 residual updates
// Prints one line per level: dofs, CG iterations, relative residual, L2
// error; then the convergence rates.
#include "gismo.h"
#include "gsXml.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

using namespace gismo;

int main(int argc, char *argv[])
{
    index_t numRefine = 5, numElevate = 1, dim = 2, gx = 0, gy = 0;
    bool last = false, full = false, mixed = false;
    std::string fn;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-r" && i + 1 < argc) numRefine = std::atoi(argv[++i]);
        else if (a == "-e" && i + 1 < argc) numElevate = std::atoi(argv[++i]);
        else if (a == "-d" && i + 1 < argc) dim = std::atoi(argv[++i]);
        else if (a == "--grid" && i + 2 < argc) { gx = std::atoi(argv[++i]); gy = std::atoi(argv[++i]); }
        else if (a == "-f" && i + 1 < argc) fn = argv[++i];
        else if (a == "--last") last = true;
        else if (a == "--full") full = true;
        else if (a == "--mixed") mixed = true;
        else { std::fprintf(stderr, "unknown argument %s\n", argv[i]); return 1; }
    }
    try {
        gsMultiPatch<> mp;
        if (!fn.empty()) {
            gsFileData<> fd(fn);
            std::printf("Loaded file %s\n", fd.lastPath().c_str());
            fd.getFirst(mp);
        } else if (gx > 0) mp = gsNurbsCreator<>::BSplineSquareGrid(gx, gy);
        else {
            mp.addPatch(dim == 3 ? gsNurbsCreator<>::BSplineCube() : gsNurbsCreator<>::BSplineSquare());
            mp.computeTopology();
        }
        const short_t d = mp.dim();
        gsFunctionExpr<> f(d == 3 ? "3*pi^2*sin(pi*x)*sin(pi*y)*sin(pi*z)" : "2*pi^2*sin(pi*x)*sin(pi*y)", d);
        gsFunctionExpr<> ms(d == 3 ? "sin(pi*x)*sin(pi*y)*sin(pi*z)" : "sin(pi*x)*sin(pi*y)", d);
        gsBoundaryConditions<> bc;
        gsFunctionExpr<> zero("0", d);
        for (auto &s : mp.boundaries())
            bc.addCondition(s.first, (boundary::side)s.second, condition_type::dirichlet, &zero);

        gsMultiBasis<> dbasis(mp, true);
        dbasis.setDegree(dbasis.maxCwiseDegree() + numElevate);
        if (last) {
            for (int r = 0; r < numRefine; ++r) dbasis.uniformRefine();
            numRefine = 0;
        }
        std::printf("Patches: %d, degree: %d\n", (int)mp.nPatches(), dbasis.minCwiseDegree());

        gsExprAssembler<> A(1, 1);
        if (mixed) A.options().setString("Precision", "mixed");
        typedef gsExprAssembler<>::geometryMap geometryMap;
        typedef gsExprAssembler<>::space space;
        typedef gsExprAssembler<>::solution solution;
        A.setIntegrationElements(dbasis);
        gsExprEvaluator<> ev(A);
        geometryMap G = A.getMap(mp);
        space u = A.getSpace(dbasis);
        auto ff = A.getCoeff(f, G);
        auto u_ex = ev.getVariable(ms, G);
        gsMatrix<> solVector;
        solution u_sol = A.getSolution(u, solVector);

        std::vector<double> l2err;
        double ma_time = 0, slv_time = 0;
        for (int r = 0; r <= numRefine; ++r) {
            dbasis.uniformRefine();
            u.setup(bc, dirichlet::l2Projection, 0);
            A.initSystem();
            auto t0 = std::chrono::steady_clock::now();
            const gsSparseMatrix<> *K;
            const gsMatrix<> *rhs;
            std::unique_ptr<gsPoissonAssembler<>> pa;
            if (full) {
                pa.reset(new gsPoissonAssembler<>(mp, dbasis, bc, f));
                if (mixed) pa->options().setString("Precision", "mixed");
                pa->assemble();
                K = &pa->matrix();
                rhs = &pa->rhs();
            } else {
                A.assemble(igrad(u, G) * igrad(u, G).tr() * meas(G), u * ff * meas(G));
                K = &A.matrix();
                rhs = &A.rhs();
            }
            auto t1 = std::chrono::steady_clock::now();
            gsConjugateGradient<> cg(*K, makeJacobiOp(*K, 1.0));
            gsOptionList o = gsConjugateGradient<>::defaultOptions();
            o.setReal("Tolerance", 1e-10);
            o.setInt("MaxIterations", 100000);
            if (mixed) o.setString("Precision", "mixed");
            cg.setOptions(o);
            solVector.setZero(K->rows(), 1);
            cg.solve(*rhs, solVector);
            auto t2 = std::chrono::steady_clock::now();
            ma_time += std::chrono::duration<double>(t1 - t0).count();
            slv_time += std::chrono::duration<double>(t2 - t1).count();
            const double e = full ? gsExprEvaluator<>(*pa).l2ErrorSinProduct(solVector)
                                  : math::sqrt(ev.integral((u_ex - u_sol).sqNorm() * meas(G)));
            l2err.push_back(e);
            std::printf("level %d dofs %d nnz %d iterations %d relres %.3e L2 %.17g\n", r, (int)K->rows(),
                        (int)K->nonZeros(), (int)cg.iterations(), cg.error(), e);
        }
        std::printf("Assembly: %.4f s  Solving: %.4f s\n", ma_time, slv_time);
        if (l2err.size() > 1) {
            std::printf("EoC (L2):");
            for (size_t i = 1; i < l2err.size(); ++i) std::printf(" %.2f", std::log(l2err[i - 1] / l2err[i]) / std::log(2.0));
            std::printf("\n");
        }
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 2;
    }
    return 0;
}
