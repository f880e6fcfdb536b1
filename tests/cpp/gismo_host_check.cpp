// Host-side checks of the C++ drop-in classes (no GPU needed): knot vector
// construction and refinement, topology detection, the geometry mapped
// into the discretisation basis, and the DOF mapper (host-only entry
// gsx_mapper_build) checked physically: two basis functions share a
// global index exactly when their control points coincide.
// Prints "name value..." lines that the Python test compares; exits
// nonzero on a failed internal check.
#include "gismo.h"

#include <cstdio>
#include <map>
#include <tuple>

using namespace gismo;

static int failures = 0;
#define CHECK(c)                                                            \
    do {                                                                    \
        if (!(c)) {                                                         \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);         \
            ++failures;                                                     \
        }                                                                   \
    } while (0)

static void print_knots(const char *name, const gsKnotVector &k)
{
    std::printf("%s %d", name, k.degree());
    for (size_t i = 0; i < k.size(); ++i) std::printf(" %.17g", k[i]);
    std::printf("\n");
}

static void print_topology(const char *name, const gsMultiPatch<> &mp)
{
    std::printf("%s_ifaces", name);
    for (auto &f : mp.interfaces())
        std::printf(" %d,%d,%d,%d,%d,%d,%d,%d", f.p1, f.s1, f.p2, f.s2, f.m0, f.f0, f.m1, f.f1);
    std::printf("\n%s_bdr", name);
    for (auto &b : mp.boundaries()) std::printf(" %d,%d", b.first, b.second);
    std::printf("\n");
}

// builds the mapper for mp refined to degree p with n elements per
// direction and checks it against control-point coincidence
static void check_mapper(const char *name, const gsMultiPatch<> &mp, int p, int refine)
{
    gsMultiBasis<> mb(mp, true);
    mb.setDegree(p);
    for (int r = 0; r < refine; ++r) mb.uniformRefine();
    const int np = mp.nPatches(), d = mp.dim();
    std::vector<int32_t> nbd(3 * np, 1);
    std::vector<std::vector<double>> ctrl(np);
    int64_t total = 0;
    for (int k = 0; k < np; ++k) {
        for (int j = 0; j < d; ++j) nbd[3 * k + j] = mb.knots(k, j).numBasis();
        ctrl[k] = detail::isoCoefs(mp.patch(k), mb, k);
        total += mb.size(k);
    }
    std::vector<int32_t> ds, ifs;
    for (auto &s : mp.boundaries()) { ds.push_back(s.first); ds.push_back(s.second); }
    for (auto &f : mp.interfaces()) {
        const int32_t r[8] = {f.p1, f.s1, f.p2, f.s2, f.m0, f.f0, f.m1, f.f1};
        ifs.insert(ifs.end(), r, r + 8);
    }
    std::vector<int32_t> dofs(total);
    int64_t nfree = 0, nglued = 0;
    const gsx_status st = gsx_mapper_build(d, np, nbd.data(), (int32_t)ds.size() / 2, ds.data(),
                                           (int32_t)ifs.size() / 8, ifs.empty() ? nullptr : ifs.data(),
                                           dofs.data(), &nfree, &nglued);
    CHECK(st == GSX_OK);
    // points (rounded to 1e-9) -> dof
    std::map<std::tuple<long long, long long, long long>, int> at;
    std::map<int, std::tuple<long long, long long, long long>> where;
    int64_t off = 0, interior_points = 0;
    bool consistent = true;
    for (int k = 0; k < np; ++k) {
        for (index_t i = 0; i < mb.size(k); ++i) {
            long long c[3] = {0, 0, 0};
            for (int r = 0; r < d; ++r) c[r] = std::llround(ctrl[k][(size_t)i * d + r] * 1e9);
            const auto key = std::make_tuple(c[0], c[1], c[2]);
            const int g = dofs[off + i];
            auto it = at.find(key);
            if (it == at.end()) at[key] = g;
            else consistent = consistent && it->second == g;
            auto jt = where.find(g);
            if (jt == where.end()) where[g] = key;
            else consistent = consistent && jt->second == key;
        }
        off += mb.size(k);
    }
    CHECK(consistent);
    CHECK((int64_t)at.size() == nglued);
    for (auto &kv : where) interior_points += kv.first < nfree;
    CHECK(interior_points == nfree);
    std::printf("%s_mapper %lld %lld %d\n", name, (long long)nfree, (long long)nglued, consistent ? 1 : 0);
}

int main()
{
    // knot vectors (gsKnotVector::initUniform, uniformRefine, degreeElevate)
    print_knots("kv_uniform", gsKnotVector(0, 1, 3, 3));
    gsKnotVector k(0, 1, 0, 2);
    k.uniformRefine();
    print_knots("kv_refine1", k);
    k.uniformRefine(2);
    print_knots("kv_refine2", k);
    gsKnotVector e(0, 1, 0, 2);
    e.degreeElevate(2);
    print_knots("kv_elevate", e);
    gsKnotVector m(0, 1, 1, 2);
    m.degreeElevate(1);
    print_knots("kv_elevate_interior", m);

    // topology of the square grid (patch index i*ny + j)
    gsMultiPatch<> grid = gsNurbsCreator<>::BSplineSquareGrid(3, 2);
    print_topology("grid3x2", grid);

    // geometry mapped into the refined degree-3 basis
    {
        gsMultiBasis<> mb(grid, true);
        mb.setDegree(3);
        for (int r = 0; r < 2; ++r) mb.uniformRefine();
        std::vector<double> c = detail::isoCoefs(grid.patch(4), mb, 4);
        std::printf("grid3x2_ctrl4");
        for (double v : c) std::printf(" %.17g", v);
        std::printf("\n");
    }
    check_mapper("grid3x2", grid, 3, 2);

    // 2D: a flipped and a rotated neighbour
    {
        gsMultiPatch<> mp;
        mp.addPatch(gsNurbsCreator<>::BSplineSquare());
        // [1,2] x [0,1] with parameter direction 0 running along -x
        gsMatrix<> C(4, 2);
        const double pts[4][2] = {{2, 0}, {1, 0}, {2, 1}, {1, 1}};
        for (int i = 0; i < 4; ++i) { C(i, 0) = pts[i][0]; C(i, 1) = pts[i][1]; }
        mp.addPatch(gsTensorBSplinePatch({gsKnotVector(0, 1, 0, 2), gsKnotVector(0, 1, 0, 2)}, C));
        // [0,1] x [1,2] with direction 0 along y and direction 1 along -x
        gsMatrix<> R(4, 2);
        const double rp[4][2] = {{1, 1}, {1, 2}, {0, 1}, {0, 2}};
        for (int i = 0; i < 4; ++i) { R(i, 0) = rp[i][0]; R(i, 1) = rp[i][1]; }
        mp.addPatch(gsTensorBSplinePatch({gsKnotVector(0, 1, 0, 2), gsKnotVector(0, 1, 0, 2)}, R));
        mp.computeTopology();
        print_topology("flip2d", mp);
        CHECK(mp.interfaces().size() == 2);
        check_mapper("flip2d", mp, 2, 2);
    }

    // 3D: a cube and a neighbour with permuted, flipped directions
    {
        gsMultiPatch<> mp;
        mp.addPatch(gsNurbsCreator<>::BSplineCube(1, 0.5, 0.5, 0.5));   // [0,1]^3
        gsMatrix<> C(8, 3);
        for (int c = 0; c < 8; ++c) {
            const int b0 = c & 1, b1 = (c >> 1) & 1, b2 = (c >> 2) & 1;
            C(c, 0) = 2 - b1;   // direction 1 along -x
            C(c, 1) = b2;       // direction 2 along y
            C(c, 2) = 1 - b0;   // direction 0 along -z
        }
        mp.addPatch(gsTensorBSplinePatch(
            {gsKnotVector(0, 1, 0, 2), gsKnotVector(0, 1, 0, 2), gsKnotVector(0, 1, 0, 2)}, C));
        mp.computeTopology();
        print_topology("perm3d", mp);
        CHECK(mp.interfaces().size() == 1);
        CHECK(mp.boundaries().size() == 10);
        check_mapper("perm3d", mp, 2, 2);
        check_mapper("perm3d_p3", mp, 3, 1);
    }
    // 2 x 2 x 1 cubes
    {
        gsMultiPatch<> mp;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) mp.addPatch(gsNurbsCreator<>::BSplineCube(1, i + 0.5, j + 0.5, 0.5));
        mp.computeTopology();
        print_topology("cubes2x2", mp);
        CHECK(mp.interfaces().size() == 4);
        check_mapper("cubes2x2", mp, 2, 2);
    }

    // multigrid hierarchy by coarsening (gsGridHierarchy::buildByCoarsening):
    // 2 x 2 grid of degree-2 squares with 8 x 8 elements each, three levels;
    // one cube of degree 3 with 4^3 elements, two levels.  Each transfer
    // prints as rows cols | rowptr | col | val.
    {
        struct Case { const char *name; int grid, dim, p, r, levels; };
        const Case cases[] = {{"hier2d", 2, 2, 2, 3, 3}, {"hier3d", 0, 3, 3, 2, 2}};
        for (const Case &cs : cases) {
            gsMultiPatch<> mp;
            if (cs.grid) mp = gsNurbsCreator<>::BSplineSquareGrid(cs.grid, cs.grid);
            else {
                mp.addPatch(gsNurbsCreator<>::BSplineCube(1, 0.5, 0.5, 0.5));
                mp.computeTopology();
            }
            gsBoundaryConditions<> bc;
            for (auto &sd : mp.boundaries()) bc.addCondition(sd.first, (boundary::side)sd.second, condition_type::dirichlet);
            gsMultiBasis<> mb(mp, true);
            mb.setDegree(cs.p);
            for (int r = 0; r < cs.r; ++r) mb.uniformRefine();
            gsOptionList opt;
            opt.addInt("Levels", "", cs.levels);
            std::vector<gsSparseMatrix<real_t, RowMajor>> T;
            std::vector<gsMultiBasis<real_t>> bases;
            gsGridHierarchy<>::buildByCoarsening(give(mb), bc, opt).moveMultiBasesTo(bases).moveTransferMatricesTo(T);
            CHECK((int)T.size() == cs.levels - 1 && (int)bases.size() == cs.levels);
            CHECK(mb.nBases() == 0);   // moved into the hierarchy
            for (size_t l = 0; l < T.size(); ++l) {
                std::printf("%s_T%d %d %d |", cs.name, (int)l, (int)T[l].rows(), (int)T[l].cols());
                for (index_t i = 0; i <= T[l].rows(); ++i) std::printf(" %d", T[l].outerIndexPtr()[i]);
                std::printf(" |");
                for (index_t k = 0; k < T[l].nonZeros(); ++k) std::printf(" %d", T[l].innerIndexPtr()[k]);
                std::printf(" |");
                for (index_t k = 0; k < T[l].nonZeros(); ++k) std::printf(" %.17g", T[l].valuePtr()[k]);
                std::printf("\n");
            }
            for (size_t l = 0; l < bases.size(); ++l) print_knots((std::string(cs.name) + "_kv" + std::to_string(l)).c_str(),
                                                                   bases[l].knots(0, 0));
        }
    }

    // source strings
    bool threw = false;
    try { gsFunctionExpr<> f("x^2", 2); } catch (const std::runtime_error &) { threw = true; }
    CHECK(threw);
    CHECK(gsFunctionExpr<>("2*pi^2*sin(pi*x)*sin(pi*y)", 2).kind() == 1);
    CHECK(gsFunctionExpr<>("sin(pi*x)*sin(pi*y)*sin(pi*z)", 3).kind() == 3);

    std::printf("failures %d\n", failures);
    return failures ? 1 : 0;
}
