// Host-side checks of gsXml.h (no GPU): the reference's own XML inputs are
// read, their interface records agree with the topology found by corner
// matching, the refined Fichera corner has the free-DOF count the
// reference's gsBasisComponents_test asserts (117, no Dirichlet sides) and
// its DOF -> topology component golden (printed for the test to compare), and
// matrices, sparse matrices and multipatches survive a write/read round trip.
// Prints "name value..." lines; exits nonzero on a failed check.
#include "gsXml.h"

#include <algorithm>
#include <cstdio>
#include <map>
#include <set>
#include <tuple>
#include <vector>

using namespace gismo;

static int failures = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                             \
        }                                                           \
    } while (0)

typedef std::tuple<int, int, int, int, int, int, int, int> Rec;

// interface as an unordered pair of (patch, side) with the tangential map
// seen from the lower patch
static std::set<Rec> canon(const gsMultiPatch<> &mp)
{
    std::set<Rec> out;
    const int d = mp.dim();
    for (auto f : mp.interfaces()) {
        if (std::make_pair(f.p1, f.s1) > std::make_pair(f.p2, f.s2)) {
            // invert the map: tangential direction j of side s2 -> direction of s1
            const int dir1 = (f.s1 - 1) / 2, dir2 = (f.s2 - 1) / 2;
            int t1[2], k = 0;
            for (int j = 0; j < d; ++j)
                if (j != dir1) t1[k++] = j;
            int m[2] = {f.m0, f.m1}, fl[2] = {f.f0, f.f1}, nm[2] = {0, 0}, nf[2] = {0, 0};
            int t2[2];
            k = 0;
            for (int j = 0; j < d; ++j)
                if (j != dir2) t2[k++] = j;
            for (int a = 0; a < d - 1; ++a)
                for (int b = 0; b < d - 1; ++b)
                    if (m[a] == t2[b]) {
                        nm[b] = t1[a];
                        nf[b] = fl[a];
                    }
            f = {f.p2, f.s2, f.p1, f.s1, nm[0], nf[0], nm[1], nf[1]};
        }
        out.insert(Rec(f.p1, f.s1, f.p2, f.s2, f.m0, f.f0, f.m1, f.f1));
    }
    return out;
}

// Free-DOF map of the multibasis of mp refined `refine` times (glued
// interfaces, all sides Dirichlet or none): dofs per patch-local function.
static int64_t mapper_dofs(const gsMultiPatch<> &mp, int refine, bool dirichlet, std::vector<int32_t> &dofs)
{
    gsMultiBasis<> mb(mp, true);
    for (int r = 0; r < refine; ++r) mb.uniformRefine();
    const int np = mp.nPatches(), d = mp.dim();
    std::vector<int32_t> nbd(3 * np, 1), ds, ifs;
    int64_t total = 0;
    for (int k = 0; k < np; ++k) {
        for (int j = 0; j < d; ++j) nbd[3 * k + j] = mb.knots(k, j).numBasis();
        total += mb.size(k);
    }
    if (dirichlet)
        for (auto &s : mp.boundaries()) { ds.push_back(s.first); ds.push_back(s.second); }
    for (auto &f : mp.interfaces()) {
        const int32_t r[8] = {f.p1, f.s1, f.p2, f.s2, f.m0, f.f0, f.m1, f.f1};
        ifs.insert(ifs.end(), r, r + 8);
    }
    dofs.assign(total, -1);
    int64_t nfree = 0, nglued = 0;
    CHECK(gsx_mapper_build(d, np, nbd.data(), (int32_t)ds.size() / 2, ds.empty() ? nullptr : ds.data(),
                           (int32_t)ifs.size() / 8, ifs.empty() ? nullptr : ifs.data(), dofs.data(), &nfree,
                           &nglued) == GSX_OK);
    return nfree;
}

// Topology components of a conforming degree-1 multipatch of boxes in the
// order of gsBoxTopology::allComponents (src/gsCore/gsBoxTopology.cpp:
// 294-337): every patch component (index j = sum_d 3^d loc_d with loc 0
// interior, 1 begin, 2 end; src/gsCore/gsBoundary.h:445-477) is keyed by
// the sorted canonical ids patch * 2^dim + corner (corner 1-based,
// src/gsCore/gsBoundary.cpp:80-108) of the corners it contains, where a
// corner's canonical representative is the coinciding corner of smallest
// (patch, corner) (getCanonicCorner); components of equal key are one;
// the list runs from dimension dim down to 0, each in ascending key order.
// Corners coincide here when their control points do.  Returns the
// component of every (patch, component index) pair and the counts per
// dimension.
static std::vector<std::vector<int>> box_components(const gsMultiPatch<> &mp, std::vector<int> &per_dim)
{
    const int np = mp.nPatches(), d = mp.dim(), nc = 1 << d;
    int ncomp = 1;
    for (int k = 0; k < d; ++k) ncomp *= 3;
    auto corner_pt = [&](int k, int c) {
        std::vector<double> x(d);
        for (int r = 0; r < d; ++r) x[r] = mp.patch(k).coefs()(c, r);
        return x;
    };
    // canonical (patch, corner) of every patch corner
    std::vector<std::pair<int, int>> canon_c(np * nc);
    for (int k = 0; k < np; ++k)
        for (int c = 0; c < nc; ++c) {
            std::pair<int, int> best(k, c + 1);
            const auto x = corner_pt(k, c);
            for (int k2 = 0; k2 < np; ++k2)
                for (int c2 = 0; c2 < nc; ++c2)
                    if (corner_pt(k2, c2) == x && std::make_pair(k2, c2 + 1) < best) best = {k2, c2 + 1};
            canon_c[k * nc + c] = best;
        }
    std::vector<std::map<std::vector<int>, int>> keys(d + 1);
    std::vector<std::vector<std::vector<int>>> key_of(np, std::vector<std::vector<int>>(ncomp));
    std::vector<std::vector<int>> dim_of(np, std::vector<int>(ncomp));
    for (int k = 0; k < np; ++k)
        for (int j = 0; j < ncomp; ++j) {
            int cdim = 0;
            for (int t = 0, jj = j; t < d; ++t, jj /= 3) cdim += (jj % 3 == 0);
            std::vector<int> key;
            for (int i = 0; i < (1 << cdim); ++i) {
                int idx = 0, ii = i, jj = j, cstride = 1;
                for (int t = 0; t < d; ++t, cstride *= 2) {
                    if (jj % 3 == 0) { idx += cstride * (ii % 2); ii /= 2; }
                    else idx += cstride * (jj % 3 - 1);
                    jj /= 3;
                }
                const auto cc = canon_c[k * nc + idx];
                key.push_back(cc.first * nc + cc.second);
            }
            std::sort(key.begin(), key.end());
            keys[cdim][key] = 0;
            key_of[k][j] = key;
            dim_of[k][j] = cdim;
        }
    int next = 0;
    per_dim.assign(d + 1, 0);
    for (int cd = d; cd >= 0; --cd)
        for (auto &kv : keys[cd]) { kv.second = next++; ++per_dim[cd]; }
    std::vector<std::vector<int>> out(np, std::vector<int>(ncomp));
    for (int k = 0; k < np; ++k)
        for (int j = 0; j < ncomp; ++j) out[k][j] = keys[dim_of[k][j]][key_of[k][j]];
    return out;
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "tests/golden/xml";
    for (const char *name : {"two_squares.xml", "ldomain.xml", "fichera.xml"}) {
        gsFileData<> fd(dir + "/" + name);
        gsMultiPatch<> mp;
        fd.getFirst(mp);
        gsMultiPatch<> mc = mp;
        mc.computeTopology();
        const bool same = canon(mp) == canon(mc) && mp.boundaries() == mc.boundaries();
        CHECK(same);
        std::printf("%s patches %d dim %d interfaces %zu boundary %zu topology_match %d\n", name,
                    (int)mp.nPatches(), mp.dim(), mp.interfaces().size(), mp.boundaries().size(), same ? 1 : 0);
    }
    {
        gsFileData<> fd(dir + "/fichera.xml");
        gsMultiPatch<> mp;
        fd.getFirst(mp);
        std::vector<int32_t> dofs;
        const int64_t n = mapper_dofs(mp, 1, false, dofs);
        std::printf("fichera_refined_free %lld\n", (long long)n);
        CHECK(n == 117);
        // component of every free DOF: after one refinement of the degree-1
        // patches each component holds exactly one function, the one at
        // local position loc_d = 0 -> begin, 1 -> interior, 2 -> end
        std::vector<int> per_dim;
        const auto comp = box_components(mp, per_dim);
        std::vector<int> of_dof(n, -1);
        bool consistent = true;
        std::printf("fichera_local_component");
        for (int k = 0; k < (int)mp.nPatches(); ++k)
            for (int l = 0; l < 27; ++l) {
                int j = 0;
                for (int t = 0, ll = l, s3 = 1; t < 3; ++t, ll /= 3, s3 *= 3) {
                    const int pos = ll % 3;
                    j += s3 * (pos == 1 ? 0 : (pos == 0 ? 1 : 2));
                }
                const int g = dofs[27 * k + l], cid = comp[k][j];
                std::printf(" %d", cid);
                if (g < 0 || g >= n) { consistent = false; continue; }
                if (of_dof[g] >= 0 && of_dof[g] != cid) consistent = false;
                of_dof[g] = cid;
            }
        std::printf("\n");
        CHECK(consistent);
        std::printf("fichera_components_per_dim");
        for (int cd = 0; cd <= 3; ++cd) std::printf(" %d", per_dim[cd]);
        std::printf("\nfichera_component_of_dof");
        for (int v : of_dof) std::printf(" %d", v);
        std::printf("\n");
        for (auto &f : mp.interfaces())
            std::printf("fichera_iface %d %d %d %d %d %d %d %d\n", f.p1, f.s1, f.p2, f.s2, f.m0, f.f0, f.m1,
                        f.f1);
    }
    {
        gsFileData<> fd(dir + "/square.xml");
        gsMultiPatch<> mp;
        fd.getFirst(mp);
        CHECK(mp.nPatches() == 1 && mp.boundaries().size() == 4 && mp.interfaces().empty());
    }
    // round trips
    {
        gsMatrix<> M(3, 2);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 2; ++j) M(i, j) = 0.1 * (i + 1) + j / 3.0;
        std::vector<index_t> outer = {0, 2, 3, 5}, inner = {0, 2, 1, 0, 2};
        std::vector<real_t> val = {4.0, -1.0 / 3.0, 2.5, -1.0 / 3.0, 1e-17};
        gsSparseMatrix<> S;
        S.setFromCompressed(3, 3, outer, inner, val);
        gsMultiPatch<> mp = gsNurbsCreator<>::BSplineSquareGrid(2, 2);
        gsFileData<> out;
        out.add(M, 7);
        out.add(S, 8);
        out.add(mp);
        out.save("/tmp/gsxml_roundtrip.xml");
        gsFileData<> in("/tmp/gsxml_roundtrip.xml");
        gsMatrix<> M2;
        gsSparseMatrix<> S2;
        gsMultiPatch<> mp2;
        in.getId(7, M2);
        in.getId(8, S2);
        in.getFirst(mp2);
        bool ok = M2.rows() == 3 && M2.cols() == 2;
        for (int i = 0; i < 3 && ok; ++i)
            for (int j = 0; j < 2; ++j) ok = ok && M2(i, j) == M(i, j);
        CHECK(ok);
        bool sok = S2.nonZeros() == 5;
        for (int k = 0; k < 5 && sok; ++k)
            sok = S2.innerIndexPtr()[k] == inner[k] && S2.valuePtr()[k] == val[k];
        for (int j = 0; j <= 3 && sok; ++j) sok = S2.outerIndexPtr()[j] == outer[j];
        CHECK(sok);
        CHECK(canon(mp2) == canon(mp) && mp2.boundaries() == mp.boundaries());
        bool cok = mp2.nPatches() == 4;
        for (int k = 0; k < 4 && cok; ++k)
            for (int i = 0; i < 4; ++i)
                for (int r = 0; r < 2; ++r) cok = cok && mp2.patch(k).coefs()(i, r) == mp.patch(k).coefs()(i, r);
        CHECK(cok);
        std::printf("roundtrip matrix %d sparse %d multipatch %d\n", ok ? 1 : 0, sok ? 1 : 0, cok ? 1 : 0);
    }
    // boundary value problem in the layout of filedata/pde/poisson2d_bvp.xml
    {
        gsFileData<> fd(dir + "/square_pair_bvp.xml");
        gsMultiPatch<> mp;
        fd.getFirst(mp);
        gsFunctionExpr<> f, ms;
        gsBoundaryConditions<> bc;
        fd.getId(1, f);
        fd.getId(2, bc);
        fd.getId(3, ms);
        CHECK(mp.nPatches() == 2 && mp.interfaces().size() == 1);
        CHECK(f.kind() == 1 && ms.kind() == 3);
        int nd = 0, nn = 0;
        for (const auto &c : bc.conditions()) {
            if (c.type == condition_type::dirichlet) {
                ++nd;
                CHECK(c.f && c.f->kind() == 3);
                CHECK(c.side == 1 || c.side == 3);
            } else {
                ++nn;
                CHECK(c.f && c.f->kind() == 4 && c.f->targetDim() == 2);
                CHECK(c.side == 2 || c.side == 4);
            }
        }
        CHECK(bc.inhomogeneous());
        std::printf("bvp dirichlet %d neumann %d\n", nd, nn);
    }
    // refused inputs
    {
        // a scalar Neumann function (the device takes fluxes)
        bool threw = false;
        try {
            gsFileData<> fd;
            fd.readString("<xml><MultiPatch parDim=\"2\" id=\"0\"><patches type=\"id_range\">1 1</patches>"
                          "<boundary name=\"n\">1 2</boundary></MultiPatch>"
                          "<Geometry type=\"TensorBSpline2\" id=\"1\"><Basis type=\"TensorBSplineBasis2\">"
                          "<Basis type=\"BSplineBasis\" index=\"0\"><KnotVector degree=\"1\">0 0 1 1</KnotVector></Basis>"
                          "<Basis type=\"BSplineBasis\" index=\"1\"><KnotVector degree=\"1\">0 0 1 1</KnotVector></Basis>"
                          "</Basis><coefs geoDim=\"2\">0 0 1 0 0 1 1 1</coefs></Geometry>"
                          "<boundaryConditions id=\"2\" multipatch=\"0\">"
                          "<Function type=\"FunctionExpr\" dim=\"2\" index=\"0\">sin(pi*x)*sin(pi*y)</Function>"
                          "<bc type=\"Neumann\" function=\"0\" name=\"n\"></bc></boundaryConditions></xml>");
            gsBoundaryConditions<> bc;
            fd.getId(2, bc);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        CHECK(threw);
    }
    {
        bool threw = false;
        try {
            gsFileData<> fd;
            fd.readString("<xml><Geometry type=\"TensorNurbs2\" id=\"0\"/></xml>");
            gsTensorBSplinePatch g;
            fd.getId(0, g);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        CHECK(threw);
        threw = false;
        try {
            gsFileData<> fd;
            fd.readString("<xml><Matrix rows=\"2\" cols=\"2\" id=\"0\">1 2 3</Matrix></xml>");
            gsMatrix<> m;
            fd.getId(0, m);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        CHECK(threw);
    }
    std::printf("failures %d\n", failures);
    return failures ? 1 : 0;
}
