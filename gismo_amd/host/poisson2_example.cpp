// poisson2_example -- the call sequence of G+Smo's examples/poisson2_example.cpp
// (gsExprAssembler + CG with diagonal preconditioning + L2 and H1 errors
// per refinement level) on the drop-in classes of gismo.h.  The domain is
// built with gsNurbsCreator instead of read from an XML file:
//   -d 2|3        unit square / unit cube (default 2)
//   --grid NX NY  NX x NY unit squares glued at their interfaces (2D)
//   -f FILE       G+Smo XML file holding the (multi)patch geometry
//   -e N          degree elevation steps (default 1: degree 2)
//   -r N          refinement loops (default 5)
//   --last        solve only on the finest level
//   --full        gsPoissonAssembler (full pattern) instead of gsExprAssembler
//   --mixed       FP32 iterations with FP64 residual updates
//   --dv N        Dirichlet values: 100 homogeneous, 101 interpolation,
//                 102 L2 projection (default: gsExprAssembler 102 as in the
//                 reference example, gsPoissonAssembler its own default 101)
//   --fd          CG preconditioned by gsPatchPreconditionersCreator::
//                 fastDiagonalizationOp (single patch) instead of Jacobi
//   --scms        CG preconditioned by gsPatchPreconditionersCreator::
//                 subspaceCorrectedMassSmootherOp (single patch, sigma 0.12)
//   --bvp         the boundary value problem of poisson2d_bvp: the domain
//                 shifted by 1/4 (u = prod sin(pi x_i) does not vanish on
//                 it), u on the west/south(/front) sides by L2 projection,
//                 the flux grad u on the others (assembleBdr); with
//                 --full u on every side
// A file with <boundaryConditions> (the layout of the reference's
// filedata/pde/poisson2d_bvp.xml: MultiPatch, source id 1, conditions id 2,
// exact solution id 3) supplies its own data.
// Prints one line per level: dofs, CG iterations, relative residual, L2
// error, H1 error (L2 + H1 seminorm, as the reference example sums them);
// then the convergence rates.
#include "gismo.h"
#include "gsXml.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

using namespace gismo;

int main(int argc, char *argv[])
{
    index_t numRefine = 5, numElevate = 1, dim = 2, gx = 0, gy = 0;
    bool last = false, full = false, mixed = false, bvp = false, fastdiag = false, scms = false;
    int dirValues = -1;   // Dirichlet values method; -1: each assembler's own default
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
        else if (a == "--bvp") bvp = true;
        else if (a == "--fd") fastdiag = true;
        else if (a == "--scms") scms = true;
        else if (a == "--dv" && i + 1 < argc) dirValues = std::atoi(argv[++i]);
        else { std::fprintf(stderr, "unknown argument %s\n", argv[i]); return 1; }
    }
    try {
        gsMultiPatch<> mp;
        std::unique_ptr<gsFileData<>> fd;
        if (!fn.empty()) {
            fd.reset(new gsFileData<>(fn));
            std::printf("Loaded file %s\n", fd->lastPath().c_str());
            fd->getFirst(mp);
        } else if (gx > 0) mp = gsNurbsCreator<>::BSplineSquareGrid(gx, gy, 1, bvp ? 0.25 : 0, bvp ? 0.25 : 0);
        else {
            const real_t o = bvp ? 0.25 : 0;
            mp.addPatch(dim == 3 ? gsNurbsCreator<>::BSplineCube(1, o + 0.5, o + 0.5, o + 0.5) : gsNurbsCreator<>::BSplineSquare(1, o, o));
            mp.computeTopology();
        }
        const short_t d = mp.dim();
        gsFunctionExpr<> f(d == 3 ? "3*pi^2*sin(pi*x)*sin(pi*y)*sin(pi*z)" : "2*pi^2*sin(pi*x)*sin(pi*y)", d);
        gsFunctionExpr<> ms(d == 3 ? "sin(pi*x)*sin(pi*y)*sin(pi*z)" : "sin(pi*x)*sin(pi*y)", d);
        gsFunctionExpr<> flux = d == 3 ? gsFunctionExpr<>("pi*cos(pi*x)*sin(pi*y)*sin(pi*z)", "pi*sin(pi*x)*cos(pi*y)*sin(pi*z)",
                                                          "pi*sin(pi*x)*sin(pi*y)*cos(pi*z)", d)
                                       : gsFunctionExpr<>("pi*cos(pi*x)*sin(pi*y)", "pi*sin(pi*x)*cos(pi*y)", d);
        gsBoundaryConditions<> bc;
        gsFunctionExpr<> zero("0", d);
        for (auto &s : mp.boundaries()) {
            const bool lower = s.second % 2 == 1;   // west, south, front
            if (!bvp) bc.addCondition(s.first, (boundary::side)s.second, condition_type::dirichlet, &zero);
            else if (lower) bc.addCondition(s.first, (boundary::side)s.second, condition_type::dirichlet, &ms);
            else bc.addCondition(s.first, (boundary::side)s.second, condition_type::neumann, &flux);
        }
        if (fd && fd->has("boundaryConditions")) {
            // the layout of filedata/pde/poisson2d_bvp.xml: source id 1,
            // conditions id 2, exact solution id 3
            fd->getId(1, f);
            fd->getId(2, bc);
            fd->getId(3, ms);
            std::printf("Boundary conditions from the file\n");
        }
        bc.setGeoMap(mp);
        // gsPoissonAssembler takes scalar Neumann data, which the device does
        // not evaluate: its --bvp variant fixes u on every side instead
        gsBoundaryConditions<> bcAll;
        for (auto &s : mp.boundaries())
            bcAll.addCondition(s.first, (boundary::side)s.second, condition_type::dirichlet, bvp ? &ms : &zero);

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

        std::vector<double> l2err, h1err;
        double ma_time = 0, slv_time = 0;
        for (int r = 0; r <= numRefine; ++r) {
            dbasis.uniformRefine();
            u.setup(bc, dirValues < 0 ? (int)dirichlet::l2Projection : dirValues, 0);
            A.initSystem();
            auto t0 = std::chrono::steady_clock::now();
            const gsSparseMatrix<> *K;
            const gsMatrix<> *rhs;
            std::unique_ptr<gsPoissonAssembler<>> pa;
            if (full) {
                pa.reset(new gsPoissonAssembler<>(mp, dbasis, bcAll, f));
                if (mixed) pa->options().setString("Precision", "mixed");
                if (dirValues >= 0) pa->options().setInt("DirichletValues", dirValues);
                pa->assemble();
                K = &pa->matrix();
                rhs = &pa->rhs();
            } else {
                A.assemble(igrad(u, G) * igrad(u, G).tr() * meas(G), u * ff * meas(G));
                // Neumann terms defined on physical space
                auto g_N = A.getBdrFunction(G);
                A.assembleBdr(bc.get("Neumann"), u * g_N.tr() * nv(G));
                K = &A.matrix();
                rhs = &A.rhs();
            }
            auto t1 = std::chrono::steady_clock::now();
            // preconditioner: diagonal, or (--fd / --scms, one patch) the fast
            // diagonalization or the subspace corrected mass smoother of
            // gsPatchPreconditionersCreator
            gsPreconditionerOp<>::Ptr prec;
            if (fastdiag)
                prec = gsPatchPreconditionersCreator<>::fastDiagonalizationOp(dbasis.basis(0), full ? bcAll : bc);
            else if (scms)
                prec = gsPatchPreconditionersCreator<>::subspaceCorrectedMassSmootherOp(dbasis.basis(0),
                                                                                       full ? bcAll : bc);
            else
                prec = gsPreconditionerOp<>::Ptr(makeJacobiOp(*K, 1.0));
            gsConjugateGradient<> cg(*K, prec);
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
            gsExprEvaluator<> evp = full ? gsExprEvaluator<>(*pa) : ev;
            const double e = math::sqrt(evp.integral((u_ex - u_sol).sqNorm() * meas(G)));
            const double h = e + math::sqrt(evp.integral((igrad(u_ex) - igrad(u_sol, G)).sqNorm() * meas(G)));
            l2err.push_back(e);
            h1err.push_back(h);
            std::printf("level %d dofs %d nnz %d iterations %d relres %.3e L2 %.17g H1 %.17g\n", r, (int)K->rows(),
                        (int)K->nonZeros(), (int)cg.iterations(), cg.error(), e, h);
        }
        std::printf("Assembly: %.4f s  Solving: %.4f s\n", ma_time, slv_time);
        if (l2err.size() > 1) {
            std::printf("EoC (L2):");
            for (size_t i = 1; i < l2err.size(); ++i) std::printf(" %.2f", std::log(l2err[i - 1] / l2err[i]) / std::log(2.0));
            std::printf("\nEoC (H1):");
            for (size_t i = 1; i < h1err.size(); ++i) std::printf(" %.2f", std::log(h1err[i - 1] / h1err[i]) / std::log(2.0));
            std::printf("\n");
        }
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 2;
    }
    return 0;
}
