// multiGrid_example -- the call sequence of G+Smo's examples/multiGrid_example.cpp
// (assemble, build the hierarchy by coarsening, gsMultiGridOp with damped
// Jacobi smoothing, CG preconditioned by one V-cycle, solveDetailed) on the
// drop-in classes of gismo.h.  The domain is built with gsNurbsCreator:
//   -d 2|3                unit square / unit cube (default 2)
//   --grid NX NY          NX x NY glued unit squares (2D)
//   -g FILE               G+Smo XML file holding the (multi)patch geometry
//   -r N                  refinement steps (default 5)
//   -p N                  spline degree (default 2)
//   -l N                  multigrid levels (default 3)
//   --MG.NumPreSmooth N   (default 1)     --MG.NumPostSmooth N (default 1)
//   --MG.Damping X        Jacobi / Richardson damping (default 0.8)
//   -s NAME               MG.Smoother: GaussSeidel (gs, default), Jacobi (j),
//                         Richardson (r), Chebyshev (c),
//                         SubspaceCorrectedMassSmoother (scms)
//   --MG.Scaling X        sigma of the scms smoother (default 0.12)
//   -t X                  tolerance (default 1e-8)
//   --Solver.MaxIterations N (default 100)
//   -i cg|d               CG preconditioned by multigrid (default) or plain
//                         multigrid iteration
//   --mixed               FP32 cycle and CG iterations, FP64 residual updates
// Prints the iteration count, the residual history and the L2 error.
#include "gismo.h"
#include "gsXml.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

using namespace gismo;

int main(int argc, char *argv[])
{
    index_t refinements = 5, degree = 2, levels = 3, pre = 1, post = 1, maxIt = 100, dim = 2, gx = 0, gy = 0;
    real_t damping = 0.8, tol = 1e-8, scaling = 0.12;
    std::string solver = "cg", geometry, smoother = "GaussSeidel";   // the reference default
    bool mixed = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
            return argv[++i];
        };
        try {
            if (a == "-r") refinements = std::atoi(next());
            else if (a == "-p") degree = std::atoi(next());
            else if (a == "-l") levels = std::atoi(next());
            else if (a == "-d") dim = std::atoi(next());
            else if (a == "-g") geometry = next();
            else if (a == "--grid") { gx = std::atoi(next()); gy = std::atoi(next()); }
            else if (a == "--MG.NumPreSmooth") pre = std::atoi(next());
            else if (a == "--MG.NumPostSmooth") post = std::atoi(next());
            else if (a == "--MG.Damping") damping = std::atof(next());
            else if (a == "-s" || a == "--MG.Smoother") smoother = next();
            else if (a == "--MG.Scaling") scaling = std::atof(next());
            else if (a == "-t") tol = std::atof(next());
            else if (a == "--Solver.MaxIterations") maxIt = std::atoi(next());
            else if (a == "-i") solver = next();
            else if (a == "--mixed") mixed = true;
            else { std::fprintf(stderr, "unknown argument %s\n", argv[i]); return 1; }
        } catch (const std::exception &ex) {
            std::fprintf(stderr, "%s\n", ex.what());
            return 1;
        }
    }
    try {
        gsMultiPatch<> mp;
        if (!geometry.empty()) {
            gsFileData<> fd(geometry);
            fd.getFirst(mp);
        } else if (gx > 0) mp = gsNurbsCreator<>::BSplineSquareGrid(gx, gy);
        else {
            mp.addPatch(dim == 3 ? gsNurbsCreator<>::BSplineCube(1, 0.5, 0.5, 0.5) : gsNurbsCreator<>::BSplineSquare());
            mp.computeTopology();
        }
        const short_t d = mp.dim();
        gsFunctionExpr<> f(d == 3 ? "3*pi^2*sin(pi*x)*sin(pi*y)*sin(pi*z)" : "2*pi^2*sin(pi*x)*sin(pi*y)", d);
        gsFunctionExpr<> zero("0", d);
        gsBoundaryConditions<> bc;
        for (auto &s : mp.boundaries())
            bc.addCondition(s.first, (boundary::side)s.second, condition_type::dirichlet, &zero);

        gsMultiBasis<> mb(mp, true);
        mb.setDegree(degree);
        for (int r = 0; r < refinements; ++r) mb.uniformRefine();

        gsPoissonAssembler<> assembler(mp, mb, bc, f, dirichlet::elimination, iFace::glue);
        if (mixed) assembler.options().setString("Precision", "mixed");
        assembler.assemble();
        std::printf("dofs %d nnz %d\n", (int)assembler.numDofs(), (int)assembler.matrix().nonZeros());

        gsOptionList mgopt;
        mgopt.addInt("Levels", "", levels);
        mgopt.addInt("NumPreSmooth", "", pre);
        mgopt.addInt("NumPostSmooth", "", post);
        mgopt.addReal("Damping", "", damping);
        mgopt.addInt("DirichletStrategy", "", dirichlet::elimination);
        mgopt.addInt("InterfaceStrategy", "", iFace::conforming);
        mgopt.addString("Precision", "", mixed ? "mixed" : "fp64");
        mgopt.addString("Smoother", "", smoother);
        mgopt.addReal("Scaling", "", scaling);

        // examples/multiGrid_example.cpp:260-327: hierarchy by coarsening,
        // gsMultiGridOp::make(matrix, transfers), Cholesky coarse solver,
        // one smoother per level.  The subspace corrected mass smoother and
        // Chebyshev come from the hierarchy the device builds from the
        // problem (make(matrix, levels, options)).
        gsMultiGridOp<>::Ptr mg;
        const bool perLevel = smoother == "Jacobi" || smoother == "j" || smoother == "Richardson" ||
                              smoother == "r" || smoother == "GaussSeidel" || smoother == "gs";
        if (perLevel) {
            std::vector<gsSparseMatrix<real_t, RowMajor>> transferMatrices;
            std::vector<gsMultiBasis<real_t>> multiBases;
            gsGridHierarchy<>::buildByCoarsening(give(mb), bc, mgopt)
                .moveMultiBasesTo(multiBases)
                .moveTransferMatricesTo(transferMatrices);
            mg = gsMultiGridOp<>::make(assembler.matrix(), transferMatrices);
            mg->setOptions(mgopt);
            mg->setCoarseSolver(makeSparseCholeskySolver(mg->matrix(0)));
            for (index_t i = 1; i < mg->numLevels(); ++i) {
                gsPreconditionerOp<>::Ptr smootherOp;
                if (smoother == "Richardson" || smoother == "r")
                    smootherOp = makeRichardsonOp(mg->matrix(i));
                else if (smoother == "Jacobi" || smoother == "j")
                    smootherOp = makeJacobiOp(mg->matrix(i));
                else
                    smootherOp = makeGaussSeidelOp(mg->matrix(i));
                smootherOp->setOptions(mgopt);
                mg->setSmoother(i, smootherOp);
            }
            std::printf("hierarchy levels %d transfers", (int)mg->numLevels());
            for (auto &t : transferMatrices) std::printf(" %dx%d", (int)t.rows(), (int)t.cols());
            std::printf("\n");
        } else {
            mg = gsMultiGridOp<>::make(assembler.matrix(), levels, mgopt);
        }

        gsOptionList sopt = gsConjugateGradient<>::defaultOptions();
        sopt.setReal("Tolerance", tol);
        sopt.setInt("MaxIterations", maxIt);
        if (mixed) sopt.setString("Precision", "mixed");

        gsMatrix<> x, errorHistory;
        x.setZero(assembler.numDofs(), 1);
        index_t iters = 0;
        real_t relres = 0;
        if (solver == "cg") {
            gsConjugateGradient<> cg(assembler.matrix(), mg);
            cg.setOptions(sopt).solveDetailed(assembler.rhs(), x, errorHistory);
            iters = cg.iterations();
            relres = cg.error();
        } else if (solver == "d") {
            // x <- x + MG(b - A x), the stationary iteration of gsMultiGridOp::step
            gsMatrixOp<> Aop(assembler.matrix());
            const real_t bn = assembler.rhs().norm();
            gsMatrix<> Ax, r(assembler.numDofs(), 1), z;
            for (iters = 0; iters < maxIt; ++iters) {
                Aop.apply(x, Ax);
                for (index_t i = 0; i < r.rows(); ++i) r(i) = assembler.rhs()(i) - Ax(i);
                relres = r.norm() / bn;
                if (relres < tol) break;
                mg->apply(r, z);
                for (index_t i = 0; i < x.rows(); ++i) x(i) += z(i);
            }
        } else {
            std::fprintf(stderr, "unknown iterative solver %s (cg | d)\n", solver.c_str());
            return 1;
        }
        std::printf("solver %s iterations %d relres %.3e\n", solver.c_str(), (int)iters, relres);
        if (errorHistory.rows() > 0) {
            std::printf("history");
            for (index_t i = 0; i < errorHistory.rows(); ++i) std::printf(" %.3e", errorHistory(i));
            std::printf("\n");
        }
        std::printf("L2 %.17g\n", gsExprEvaluator<>(assembler).l2ErrorSinProduct(x));
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 2;
    }
    return 0;
}
