// gismo.h -- drop-in C++ interface for G+Smo's isogeometric Poisson path,
// backed by libgsx (HIP, gfx950) through its C ABI (include/gsx.h).
//
// Mirrors the names, signatures and call sequences of the reference for the
// path it replaces (paths relative to the G+Smo source tree):
//   containers    gsMatrix, gsSparseMatrix (column-major, int indices)
//                 src/gsMatrix/gsMatrix.h:38, src/gsMatrix/gsSparseMatrix.h:139-448
//   geometry      gsKnotVector, gsTensorBSpline, gsMultiPatch, gsMultiBasis,
//                 gsNurbsCreator::BSplineSquare/BSplineSquareGrid/BSplineCube
//                 src/gsNurbs/gsKnotVector.h, src/gsCore/gsMultiPatch.h,
//                 src/gsNurbs/gsNurbsCreator.hpp:655-745
//   conditions    gsBoundaryConditions (homogeneous Dirichlet sides)
//                 src/gsPde/gsBoundaryConditions.h:25-981
//   assembly      gsExprAssembler (igrad(u,G)*igrad(u,G).tr()*meas(G) [*kappa],
//                 u*ff*meas(G)) src/gsAssembler/gsExprAssembler.h:29-833;
//                 gsPoissonAssembler src/gsAssembler/gsPoissonAssembler.h:35-142
//   solvers       gsLinearOperator, gsMatrixOp, gsJacobiOp/makeJacobiOp,
//                 gsIterativeSolver, gsConjugateGradient, gsMultiGridOp
//                 src/gsSolver/*.h, src/gsMultiGrid/gsMultiGrid.h
//   evaluation    gsExprEvaluator::integral of the L2 error against the
//                 manufactured solution src/gsAssembler/gsExprEvaluator.h:382-450
//
// Restrictions of this drop-in (each raises std::runtime_error when hit):
// isoparametric patches, or multilinear single-element geometries (the
// gsNurbsCreator boxes) mapped into the discretisation basis; equal degree
// in all directions, degree 1..4; homogeneous Dirichlet conditions; the
// source d*pi^2*prod sin(pi x_i) (as its gsFunctionExpr string) or a
// per-element constant.
#pragma once

#include <gsx.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <cstdlib>
#include <string>
#include <vector>

namespace gismo {

typedef double real_t;
typedef int index_t;
typedef short short_t;

/// the reference's info stream (src/gsCore/gsDebug.h)
inline std::ostream &gsInfo = std::cout;

// ---------------------------------------------------------------- errors
namespace detail {

inline gsx_ctx *ctx()
{
    struct Holder {
        gsx_ctx *h = nullptr;
        Holder()
        {
            if (gsx_ctx_create(0, &h) != GSX_OK) h = nullptr;
        }
        ~Holder()
        {
            if (h) gsx_ctx_destroy(h);
        }
    };
    static Holder holder;
    if (!holder.h) throw std::runtime_error("gismo (libgsx): no GPU device available");
    return holder.h;
}

inline void check(gsx_status s, const char *what)
{
    if (s == GSX_OK) return;
    std::ostringstream os;
    os << what << ": " << gsx_last_error(ctx());
    if (s == GSX_ERR_ARG || s == GSX_ERR_JACOBIAN) throw std::runtime_error(os.str());
    throw std::runtime_error(os.str());
}

struct DeviceVec {
    void *p = nullptr;
    int64_t bytes = 0;
    explicit DeviceVec(int64_t b) : bytes(b) { check(gsx_device_alloc(ctx(), b, &p), "device allocation"); }
    ~DeviceVec() { if (p) gsx_device_free(ctx(), p); }
    DeviceVec(const DeviceVec &) = delete;
    DeviceVec &operator=(const DeviceVec &) = delete;
    double *d() const { return static_cast<double *>(p); }
};

}  // namespace detail

// ---------------------------------------------------------------- gsMatrix
/// Dense column-major matrix (the subset of gsMatrix the path uses).
template <class T = real_t>
class gsMatrix {
public:
    gsMatrix() = default;
    gsMatrix(index_t r, index_t c) { resize(r, c); }
    index_t rows() const { return m_rows; }
    index_t cols() const { return m_cols; }
    index_t size() const { return m_rows * m_cols; }
    void resize(index_t r, index_t c)
    {
        m_rows = r;
        m_cols = c;
        m_data.assign((size_t)r * c, T(0));
    }
    void setZero() { std::fill(m_data.begin(), m_data.end(), T(0)); }
    void setZero(index_t r, index_t c) { resize(r, c); }
    T &operator()(index_t i, index_t j) { return m_data[(size_t)j * m_rows + i]; }
    const T &operator()(index_t i, index_t j) const { return m_data[(size_t)j * m_rows + i]; }
    T &operator()(index_t i) { return m_data[i]; }
    const T &operator()(index_t i) const { return m_data[i]; }
    T &at(index_t i) { return m_data.at(i); }
    const T &at(index_t i) const { return m_data.at(i); }
    T *data() { return m_data.data(); }
    const T *data() const { return m_data.data(); }
    T norm() const
    {
        T s = 0;
        for (T v : m_data) s += v * v;
        return std::sqrt(s);
    }
    static gsMatrix Zero(index_t r, index_t c) { return gsMatrix(r, c); }
    /// uniform values in [-1, 1] from std::rand (Eigen's setRandom)
    void setRandom(index_t r, index_t c)
    {
        resize(r, c);
        for (T &v : m_data) v = T(2) * (T)std::rand() / (T)RAND_MAX - T(1);
    }

private:
    index_t m_rows = 0, m_cols = 0;
    std::vector<T> m_data;
};

template <class T = real_t>
using gsVector = gsMatrix<T>;

// ---------------------------------------------------------------- gsSparseMatrix
/// Storage orders of gsSparseMatrix (Eigen's ColMajor / RowMajor flags).
enum { ColMajor = 0, RowMajor = 1 };

/// Compressed sparse matrix with int indices (Eigen's compressed storage
/// accessors), column-major by default, row-major with RowMajor (the
/// multigrid transfers, gsSparseMatrix<T, RowMajor>).  Matrices produced by
/// the assemblers are symmetric bitwise and keep a handle to their device
/// copy, which the solvers use without a host round trip.
template <class T = real_t, int _Options = ColMajor>
class gsSparseMatrix {
public:
    struct Triplet {
        index_t row, col;
        T value;
    };
    gsSparseMatrix() = default;
    gsSparseMatrix(index_t r, index_t c) : m_rows(r), m_cols(c), m_outer((_Options == RowMajor ? r : c) + 1, 0) {}
    index_t rows() const { return m_rows; }
    index_t cols() const { return m_cols; }
    index_t outerSize() const { return _Options == RowMajor ? m_rows : m_cols; }
    index_t nonZeros() const { return (index_t)m_val.size(); }
    bool isCompressed() const { return true; }
    const index_t *outerIndexPtr() const { return m_outer.data(); }
    const index_t *innerIndexPtr() const { return m_inner.data(); }
    const T *valuePtr() const { return m_val.data(); }
    T coeff(index_t i, index_t j) const
    {
        const index_t o = _Options == RowMajor ? i : j, in = _Options == RowMajor ? j : i;
        auto b = m_inner.begin() + m_outer[o], e = m_inner.begin() + m_outer[o + 1];
        auto it = std::lower_bound(b, e, in);
        return (it != e && *it == in) ? m_val[it - m_inner.begin()] : T(0);
    }
    T at(index_t i, index_t j) const { return coeff(i, j); }
    /// Fill from compressed outer vectors (takes the arrays).
    void setFromCompressed(index_t r, index_t c, std::vector<index_t> outer, std::vector<index_t> inner,
                           std::vector<T> val)
    {
        m_rows = r;
        m_cols = c;
        m_outer = std::move(outer);
        m_inner = std::move(inner);
        m_val = std::move(val);
        m_dev.reset();
    }
    /// setFromTriplets with take_first (the first entry of a duplicated
    /// position is kept, src/gsCore/gsMultiBasis.hpp:175-176), or summing
    /// duplicates when takeFirst is false (Eigen's default).
    void setFromTriplets(index_t r, index_t c, const std::vector<Triplet> &t, bool takeFirst)
    {
        m_rows = r;
        m_cols = c;
        const index_t no = outerSize();
        std::vector<std::vector<std::pair<index_t, T>>> v((size_t)no);
        for (const Triplet &e : t) {
            const index_t o = _Options == RowMajor ? e.row : e.col, in = _Options == RowMajor ? e.col : e.row;
            auto &vec = v[(size_t)o];
            auto it = std::find_if(vec.begin(), vec.end(), [&](const std::pair<index_t, T> &q) { return q.first == in; });
            if (it == vec.end()) vec.push_back({in, e.value});
            else if (!takeFirst) it->second += e.value;
        }
        m_outer.assign(no + 1, 0);
        m_inner.clear();
        m_val.clear();
        for (index_t o = 0; o < no; ++o) {
            auto &vec = v[(size_t)o];
            std::sort(vec.begin(), vec.end(),
                      [](const std::pair<index_t, T> &x, const std::pair<index_t, T> &y) { return x.first < y.first; });
            for (auto &q : vec) {
                m_inner.push_back(q.first);
                m_val.push_back(q.second);
            }
            m_outer[o + 1] = (index_t)m_inner.size();
        }
        m_dev.reset();
    }
    /// y = A x on the host
    gsMatrix<T> operator*(const gsMatrix<T> &x) const
    {
        gsMatrix<T> y(m_rows, 1);
        for (index_t o = 0; o < outerSize(); ++o)
            for (index_t k = m_outer[o]; k < m_outer[o + 1]; ++k) {
                if (_Options == RowMajor) y(o) += m_val[k] * x(m_inner[k]);
                else y(m_inner[k]) += m_val[k] * x(o);
            }
        return y;
    }
    /// the transpose: the same arrays read in the other storage order
    gsSparseMatrix<T, 1 - _Options> transpose() const
    {
        gsSparseMatrix<T, 1 - _Options> t;
        t.setFromCompressed(m_cols, m_rows, m_outer, m_inner, m_val);
        return t;
    }
    /// row-major CSR arrays of this matrix (for a ColMajor matrix the
    /// compressed columns are transposed)
    void toCsr(std::vector<int32_t> &rp, std::vector<int32_t> &ci, std::vector<double> &v) const
    {
        if (_Options == RowMajor) {
            rp.assign(m_outer.begin(), m_outer.end());
            ci.assign(m_inner.begin(), m_inner.end());
            v.assign(m_val.begin(), m_val.end());
            return;
        }
        rp.assign(m_rows + 1, 0);
        ci.resize(m_inner.size());
        v.resize(m_val.size());
        for (size_t k = 0; k < m_inner.size(); ++k) rp[m_inner[k] + 1]++;
        for (index_t i = 0; i < m_rows; ++i) rp[i + 1] += rp[i];
        std::vector<int32_t> pos(rp.begin(), rp.end() - 1);
        for (index_t j = 0; j < m_cols; ++j)
            for (index_t k = m_outer[j]; k < m_outer[j + 1]; ++k) {
                const int r = m_inner[k];
                ci[pos[r]] = j;
                v[pos[r]++] = m_val[k];
            }
    }
    // device twin (set by the assemblers)
    std::shared_ptr<gsx_matrix> devicePtr() const { return m_dev; }
    void setDevice(std::shared_ptr<gsx_matrix> d) { m_dev = std::move(d); }
    // the problem it was assembled on (multigrid needs the discretisation)
    std::shared_ptr<gsx_problem> problemPtr() const { return m_prob; }
    void setProblem(std::shared_ptr<gsx_problem> p) { m_prob = std::move(p); }

private:
    index_t m_rows = 0, m_cols = 0;
    std::vector<index_t> m_outer, m_inner;
    std::vector<T> m_val;
    std::shared_ptr<gsx_matrix> m_dev;
    std::shared_ptr<gsx_problem> m_prob;
};

/// give(x): move semantics of the reference's give() (src/gsCore/gsMemory.h)
template <class X>
X give(X &x) { return std::move(x); }

// ---------------------------------------------------------------- options
/// Typed key/value options (gsOptionList, src/gsIO/gsOptionList.h:32).
class gsOptionList {
public:
    gsOptionList &addInt(const std::string &k, const std::string & = "", index_t v = 0) { m_i[k] = v; return *this; }
    gsOptionList &addReal(const std::string &k, const std::string & = "", real_t v = 0) { m_r[k] = v; return *this; }
    gsOptionList &addSwitch(const std::string &k, const std::string & = "", bool v = false) { m_b[k] = v; return *this; }
    gsOptionList &addString(const std::string &k, const std::string & = "", const std::string &v = "") { m_s[k] = v; return *this; }
    void setInt(const std::string &k, index_t v) { m_i[k] = v; }
    void setReal(const std::string &k, real_t v) { m_r[k] = v; }
    void setSwitch(const std::string &k, bool v) { m_b[k] = v; }
    void setString(const std::string &k, const std::string &v) { m_s[k] = v; }
    index_t askInt(const std::string &k, index_t def = 0) const { auto it = m_i.find(k); return it == m_i.end() ? def : it->second; }
    real_t askReal(const std::string &k, real_t def = 0) const { auto it = m_r.find(k); return it == m_r.end() ? def : it->second; }
    bool askSwitch(const std::string &k, bool def = false) const { auto it = m_b.find(k); return it == m_b.end() ? def : it->second; }
    std::string askString(const std::string &k, const std::string &def = "") const { auto it = m_s.find(k); return it == m_s.end() ? def : it->second; }
    index_t getInt(const std::string &k) const { return m_i.at(k); }
    real_t getReal(const std::string &k) const { return m_r.at(k); }

private:
    std::map<std::string, index_t> m_i;
    std::map<std::string, real_t> m_r;
    std::map<std::string, bool> m_b;
    std::map<std::string, std::string> m_s;
};

// ---------------------------------------------------------------- boundary
namespace boundary {
enum side { none = 0, west = 1, east = 2, south = 3, north = 4, front = 5, back = 6 };
}
namespace condition_type {
enum type { dirichlet = 0, neumann = 1, robin = 2 };
}
namespace dirichlet {
// values of gsAssemblerOptions.h:26-45
enum strategy { elimination = 11, nitsche = 12 };
enum values { homogeneous = 100, interpolation = 101, l2Projection = 102, user = 103 };
}
namespace iFace {
enum strategy { glue = 1, dg = 2, conforming = 1 };
}

// ---------------------------------------------------------------- knots
/// Knot vector; uniform construction follows gsKnotVector::initUniform
/// (src/gsNurbs/gsKnotVector.hpp:643-688): interior knot i = first + i*h.
class gsKnotVector {
public:
    gsKnotVector() = default;
    gsKnotVector(real_t first, real_t last, unsigned interior, unsigned mult_ends, unsigned mult_interior = 1,
                 int degree = -1)
    {
        m_deg = degree == -1 ? (int)mult_ends - 1 : degree;
        const real_t h = (last - first) / (real_t)(interior + 1);
        m_k.assign(mult_ends, first);
        for (unsigned i = 1; i <= interior; ++i) m_k.insert(m_k.end(), mult_interior, first + (real_t)i * h);
        m_k.insert(m_k.end(), mult_ends, last);
    }
    gsKnotVector(std::vector<real_t> knots, int degree) : m_k(std::move(knots)), m_deg(degree) {}
    int degree() const { return m_deg; }
    size_t size() const { return m_k.size(); }
    real_t operator[](size_t i) const { return m_k[i]; }
    const std::vector<real_t> &knots() const { return m_k; }
    index_t numBasis() const { return (index_t)m_k.size() - m_deg - 1; }
    /// insert numKnots uniformly spaced knots in every non-empty span
    /// (gsKnotVector::uniformRefine / getUniformRefinementKnots,
    /// src/gsNurbs/gsKnotVector.hpp:820-834,1048-1062)
    void uniformRefine(int numKnots = 1)
    {
        std::vector<real_t> out;
        for (size_t i = 0; i + 1 < m_k.size(); ++i) {
            out.push_back(m_k[i]);
            if (m_k[i + 1] > m_k[i]) {
                const real_t step = (m_k[i + 1] - m_k[i]) / (real_t)(numKnots + 1);
                for (int j = 1; j <= numKnots; ++j) out.push_back(m_k[i] + (real_t)j * step);
            }
        }
        out.push_back(m_k.back());
        m_k.swap(out);
    }
    /// raise the degree by r, keeping the continuity: every distinct knot
    /// gains r in multiplicity (gsBSplineBasis::degreeElevate)
    void degreeElevate(int r = 1)
    {
        std::vector<real_t> out;
        for (size_t i = 0; i < m_k.size(); ++i) {
            out.push_back(m_k[i]);
            if (i + 1 == m_k.size() || m_k[i + 1] != m_k[i]) out.insert(out.end(), r, m_k[i]);
        }
        m_k.swap(out);
        m_deg += r;
    }

private:
    std::vector<real_t> m_k;
    int m_deg = 0;
};

// ---------------------------------------------------------------- patches
/// Tensor-product B-spline patch: knot vectors and control points
/// (rows = basis functions, lexicographic with direction 0 fastest).
class gsTensorBSplinePatch {
public:
    gsTensorBSplinePatch() = default;
    gsTensorBSplinePatch(std::vector<gsKnotVector> kv, gsMatrix<real_t> coefs)
        : m_kv(std::move(kv)), m_coefs(std::move(coefs)) {}
    short_t domainDim() const { return (short_t)m_kv.size(); }
    short_t targetDim() const { return (short_t)m_coefs.cols(); }
    const gsKnotVector &knots(int k) const { return m_kv[k]; }
    const gsMatrix<real_t> &coefs() const { return m_coefs; }
    gsMatrix<real_t> &coefs() { return m_coefs; }
    index_t basisSize() const
    {
        index_t n = 1;
        for (auto &k : m_kv) n *= k.numBasis();
        return n;
    }
    /// x = G(xi) of a multilinear single-element patch, as nested linear
    /// interpolation along direction 0, then 1, then 2 (a box [a, b]
    /// gives a + t (b - a) in every coordinate)
    void evalMultilinear(const real_t *xi, real_t *x) const
    {
        const int d = domainDim(), D = targetDim();
        for (int r = 0; r < D; ++r) {
            real_t v[8];
            for (int c = 0; c < (1 << d); ++c) v[c] = m_coefs(c, r);
            for (int k = 0; k < d; ++k) {
                const real_t t = (xi[k] - m_kv[k][0]) / (m_kv[k][m_kv[k].size() - 1] - m_kv[k][0]);
                for (int c = 0; c < (1 << (d - k - 1)); ++c) v[c] = v[2 * c] + t * (v[2 * c + 1] - v[2 * c]);
            }
            x[r] = v[0];
        }
    }
    bool isMultilinearBox() const
    {
        for (auto &k : m_kv)
            if (k.degree() != 1 || k.numBasis() != 2) return false;
        return true;
    }

private:
    std::vector<gsKnotVector> m_kv;
    gsMatrix<real_t> m_coefs;
};

template <short_t d, class T = real_t>
using gsTensorBSpline = gsTensorBSplinePatch;

/// Patches plus their topology (conforming interfaces, boundary sides).
template <class T = real_t>
class gsMultiPatch {
public:
    struct Interface { int p1, s1, p2, s2, m0, f0, m1, f1; };
    void addPatch(const gsTensorBSplinePatch &p) { m_p.push_back(p); }
    void addPatch(std::unique_ptr<gsTensorBSplinePatch> p) { m_p.push_back(*p); }
    index_t nPatches() const { return (index_t)m_p.size(); }
    const gsTensorBSplinePatch &patch(index_t i) const { return m_p[i]; }
    short_t dim() const { return m_p.empty() ? 0 : m_p[0].domainDim(); }
    /// find conforming interfaces by matching side corners
    /// (gsMultiPatch::computeTopology, src/gsCore/gsMultiPatch.hpp:366-493),
    /// for the multilinear box patches and isoparametric patches whose side
    /// corners are control points
    bool computeTopology(real_t tol = 1e-4)
    {
        m_if.clear();
        const int d = dim();
        auto side_corners = [&](int k, int side, std::vector<std::vector<real_t>> &cs,
                                std::vector<std::vector<int>> &idx) {
            const gsTensorBSplinePatch &P = m_p[k];
            int nb[3] = {1, 1, 1};
            for (int j = 0; j < d; ++j) nb[j] = P.knots(j).numBasis();
            const int dir = (side - 1) / 2, par = (side - 1) % 2;
            cs.clear();
            idx.clear();
            for (int c = 0; c < (1 << (d - 1)); ++c) {
                int ii[3] = {0, 0, 0}, t = 0;
                for (int j = 0; j < d; ++j) {
                    if (j == dir) { ii[j] = par ? nb[j] - 1 : 0; continue; }
                    ii[j] = (c >> t) & 1 ? nb[j] - 1 : 0;
                    ++t;
                }
                const int l = ii[0] + nb[0] * (ii[1] + nb[1] * ii[2]);
                std::vector<real_t> x(P.targetDim());
                for (int r = 0; r < P.targetDim(); ++r) x[r] = P.coefs()(l, r);
                cs.push_back(x);
                idx.push_back({ii[0], ii[1], ii[2]});
            }
        };
        std::vector<std::vector<int>> used(m_p.size(), std::vector<int>(2 * d + 1, 0));
        for (int a = 0; a < (int)m_p.size(); ++a)
            for (int sa = 1; sa <= 2 * d; ++sa) {
                if (used[a][sa]) continue;
                std::vector<std::vector<real_t>> ca, cb;
                std::vector<std::vector<int>> ia, ib;
                side_corners(a, sa, ca, ia);
                for (int b = a + 1; b < (int)m_p.size() && !used[a][sa]; ++b)
                    for (int sb = 1; sb <= 2 * d && !used[a][sa]; ++sb) {
                        if (used[b][sb]) continue;
                        side_corners(b, sb, cb, ib);
                        // corner correspondence
                        std::vector<int> match(ca.size(), -1);
                        bool ok = true;
                        for (size_t i = 0; i < ca.size() && ok; ++i) {
                            for (size_t j = 0; j < cb.size(); ++j) {
                                real_t dist = 0;
                                for (size_t r = 0; r < ca[i].size(); ++r) dist += std::fabs(ca[i][r] - cb[j][r]);
                                if (dist < tol) { match[i] = (int)j; break; }
                            }
                            ok = match[i] >= 0;
                        }
                        if (!ok) continue;
                        // tangential directions of a -> directions/orientation of b
                        const int dira = (sa - 1) / 2, dirb = (sb - 1) / 2;
                        int rec[8] = {a, sa, b, sb, 0, 0, 0, 0}, t = 0;
                        for (int j = 0; j < d; ++j) {
                            if (j == dira) continue;
                            // corner 0 vs the corner that moves along j in a
                            const int c1 = 1 << t;
                            const std::vector<int> &p0 = ib[match[0]], &p1 = ib[match[c1]];
                            int mj = -1, fj = 0;
                            for (int m = 0; m < d; ++m)
                                if (m != dirb && p0[m] != p1[m]) { mj = m; fj = p1[m] < p0[m]; }
                            rec[4 + 2 * t] = mj;
                            rec[5 + 2 * t] = fj;
                            ++t;
                        }
                        m_if.push_back({rec[0], rec[1], rec[2], rec[3], rec[4], rec[5], rec[6], rec[7]});
                        used[a][sa] = used[b][sb] = 1;
                    }
            }
        m_bdr.clear();
        for (int a = 0; a < (int)m_p.size(); ++a)
            for (int s = 1; s <= 2 * d; ++s)
                if (!used[a][s]) m_bdr.push_back({a, s});
        return true;
    }
    const std::vector<Interface> &interfaces() const { return m_if; }
    const std::vector<std::pair<int, int>> &boundaries() const { return m_bdr; }
    /// boundary sides listed under a name (e.g. read from an XML file)
    const std::map<std::string, std::vector<std::pair<int, int>>> &namedBoundaries() const { return m_named; }
    /// topology given explicitly (as in an XML MultiPatch); sides neither in
    /// an interface nor listed are added to the boundary
    void setTopology(std::vector<Interface> ifs, std::vector<std::pair<int, int>> bdr,
                     std::map<std::string, std::vector<std::pair<int, int>>> named = {})
    {
        m_if = std::move(ifs);
        m_named = std::move(named);
        const int d = dim();
        std::vector<std::vector<int>> used(m_p.size(), std::vector<int>(2 * d + 1, 0));
        for (auto &f : m_if) used.at(f.p1).at(f.s1) = used.at(f.p2).at(f.s2) = 1;
        for (auto &b : bdr) used.at(b.first).at(b.second) = 2;
        m_bdr = std::move(bdr);
        for (int k = 0; k < (int)m_p.size(); ++k)
            for (int s = 1; s <= 2 * d; ++s)
                if (!used[k][s]) m_bdr.push_back({k, s});
        std::sort(m_bdr.begin(), m_bdr.end());
        m_bdr.erase(std::unique(m_bdr.begin(), m_bdr.end()), m_bdr.end());
    }

private:
    std::vector<gsTensorBSplinePatch> m_p;
    std::vector<Interface> m_if;
    std::vector<std::pair<int, int>> m_bdr;
    std::map<std::string, std::vector<std::pair<int, int>>> m_named;
};

/// Discretisation bases, one tensor B-spline basis per patch.
/// gsTensorBSplineBasis (src/gsTensor/gsTensorBSplineBasis.h): the knot
/// vectors of one patch's tensor basis
template <class T = real_t>
class gsTensorBSplineBasis {
public:
    explicit gsTensorBSplineBasis(std::vector<gsKnotVector> kv) : m_kv(std::move(kv)) {}
    short_t dim() const { return (short_t)m_kv.size(); }
    const gsKnotVector &knots(int dir) const { return m_kv[dir]; }
    int degree(int dir) const { return m_kv[dir].degree(); }
    index_t size() const
    {
        index_t n = 1;
        for (auto &k : m_kv) n *= k.numBasis();
        return n;
    }

private:
    std::vector<gsKnotVector> m_kv;
};

template <class T = real_t>
class gsMultiBasis {
public:
    gsMultiBasis() = default;
    explicit gsMultiBasis(const gsMultiPatch<T> &mp, bool = false)
    {
        for (index_t k = 0; k < mp.nPatches(); ++k) {
            std::vector<gsKnotVector> kv;
            for (int j = 0; j < mp.patch(k).domainDim(); ++j) kv.push_back(mp.patch(k).knots(j));
            m_kv.push_back(kv);
        }
        m_if = mp.interfaces();   // the box topology the reference's gsMultiBasis keeps
    }
    const std::vector<typename gsMultiPatch<T>::Interface> &interfaces() const { return m_if; }
    /// remove every other interior knot of every direction (gsKnotVector::
    /// coarsen(1, 1), src/gsNurbs/gsKnotVector.hpp:951-985: the distinct
    /// interior knots number 1, 3, 5, ... are removed)
    void uniformCoarsen(int numKnots = 1)
    {
        if (numKnots != 1) throw std::runtime_error("gsMultiBasis::uniformCoarsen: one knot per span only");
        for (auto &p : m_kv)
            for (auto &k : p) {
                const std::vector<real_t> &U = k.knots();
                std::vector<real_t> uniq;
                for (real_t u : U)
                    if (uniq.empty() || u != uniq.back()) uniq.push_back(u);
                std::vector<real_t> removed;
                for (size_t i = 1; i + 1 < uniq.size(); i += 2) removed.push_back(uniq[i]);
                std::vector<real_t> out;
                for (real_t u : U)
                    if (!std::binary_search(removed.begin(), removed.end(), u)) out.push_back(u);
                k = gsKnotVector(out, k.degree());
            }
    }
    index_t nBases() const { return (index_t)m_kv.size(); }
    void degreeElevate(int r = 1)
    {
        for (auto &p : m_kv)
            for (auto &k : p) k.degreeElevate(r);
    }
    void setDegree(int deg)
    {
        for (auto &p : m_kv)
            for (auto &k : p)
                if (k.degree() < deg) k.degreeElevate(deg - k.degree());
    }
    void uniformRefine(int numKnots = 1)
    {
        for (auto &p : m_kv)
            for (auto &k : p) k.uniformRefine(numKnots);
    }
    const gsKnotVector &knots(int patch, int dir) const { return m_kv[patch][dir]; }
    int degree(int patch = 0, int dir = 0) const { return m_kv[patch][dir].degree(); }
    int maxCwiseDegree() const
    {
        int m = 0;
        for (auto &p : m_kv)
            for (auto &k : p) m = std::max(m, k.degree());
        return m;
    }
    int minCwiseDegree() const
    {
        int m = 1 << 20;
        for (auto &p : m_kv)
            for (auto &k : p) m = std::min(m, k.degree());
        return m;
    }
    index_t size(int patch) const
    {
        index_t n = 1;
        for (auto &k : m_kv[patch]) n *= k.numBasis();
        return n;
    }
    index_t totalSize() const
    {
        index_t n = 0;
        for (int k = 0; k < nBases(); ++k) n += size(k);
        return n;
    }
    /// the tensor B-spline basis of one patch (gsMultiBasis::basis / operator[])
    gsTensorBSplineBasis<T> basis(int patch) const { return gsTensorBSplineBasis<T>(m_kv[patch]); }
    gsTensorBSplineBasis<T> operator[](int patch) const { return basis(patch); }

private:
    std::vector<std::vector<gsKnotVector>> m_kv;
    std::vector<typename gsMultiPatch<T>::Interface> m_if;
};

/// gsNurbsCreator subset (src/gsNurbs/gsNurbsCreator.hpp:655-745): degree-1
/// single-element boxes.
template <class T = real_t>
struct gsNurbsCreator {
    static std::unique_ptr<gsTensorBSplinePatch> BSplineSquare(real_t r = 1, real_t x = 0, real_t y = 0)
    {
        gsMatrix<real_t> C(4, 2);
        const real_t pts[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
        for (int i = 0; i < 4; ++i) {
            C(i, 0) = r * pts[i][0] + x;
            C(i, 1) = r * pts[i][1] + y;
        }
        return std::unique_ptr<gsTensorBSplinePatch>(
            new gsTensorBSplinePatch({gsKnotVector(0, 1, 0, 2), gsKnotVector(0, 1, 0, 2)}, C));
    }
    /// n x m squares; patch index i*m + j at [lx + r i, ..] x [ly + r j, ..]
    static gsMultiPatch<T> BSplineSquareGrid(int n, int m, real_t r = 1, real_t lx = 0, real_t ly = 0)
    {
        gsMultiPatch<T> mp;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < m; ++j) mp.addPatch(BSplineSquare(r, lx + r * (real_t)i, ly + r * (real_t)j));
        mp.computeTopology();
        return mp;
    }
    /// cube of side r centred at (x, y, z) (src/gsNurbs/gsNurbsCreator.hpp:731-745)
    static std::unique_ptr<gsTensorBSplinePatch> BSplineCube(real_t r = 1, real_t x = 0, real_t y = 0, real_t z = 0)
    {
        gsMatrix<real_t> C(8, 3);
        for (int i = 0; i < 8; ++i) {
            C(i, 0) = r * ((real_t)(i & 1) - 0.5) + x;
            C(i, 1) = r * ((real_t)((i >> 1) & 1) - 0.5) + y;
            C(i, 2) = r * ((real_t)((i >> 2) & 1) - 0.5) + z;
        }
        return std::unique_ptr<gsTensorBSplinePatch>(new gsTensorBSplinePatch(
            {gsKnotVector(0, 1, 0, 2), gsKnotVector(0, 1, 0, 2), gsKnotVector(0, 1, 0, 2)}, C));
    }
};

// ---------------------------------------------------------------- functions
/// Source functions the device path evaluates itself.
template <class T = real_t>
class gsFunctionExpr {
public:
    gsFunctionExpr() = default;
    /// Scalar functions evaluated on the device: "0"; the manufactured
    /// solution prod sin(pi x_i) and its source d*pi^2*prod sin(pi x_i); an
    /// affine function c + a*x + b*y (+ e*z) (blanks ignored)
    gsFunctionExpr(const std::string &expr, short_t dim) : m_dim(dim)
    {
        const std::string e = strip(expr);
        if (e == "0") { m_kind = 2; return; }
        const std::string u2 = "sin(pi*x)*sin(pi*y)", u3 = "sin(pi*x)*sin(pi*y)*sin(pi*z)";
        if ((dim == 2 && e == "2*pi^2*" + u2) || (dim == 3 && e == "3*pi^2*" + u3)) { m_kind = 1; return; }
        if ((dim == 2 && e == u2) || (dim == 3 && e == u3)) { m_kind = 3; return; }
        if (affine(e, dim, m_c)) { m_kind = 5; return; }
        throw std::runtime_error("gsFunctionExpr: only 0, affine functions, the manufactured solution "
                                 "prod sin(pi x_i) and its source d*pi^2*prod sin(pi x_i) are evaluated on the "
                                 "device, got '" + expr + "'");
    }
    /// Vector functions (Neumann fluxes): the gradient of the manufactured
    /// solution, or constant components
    gsFunctionExpr(const std::string &c0, const std::string &c1, short_t dim) : gsFunctionExpr(std::vector<std::string>{c0, c1}, dim) {}
    gsFunctionExpr(const std::string &c0, const std::string &c1, const std::string &c2, short_t dim)
        : gsFunctionExpr(std::vector<std::string>{c0, c1, c2}, dim) {}
    gsFunctionExpr(const std::vector<std::string> &comp, short_t dim) : m_dim(dim), m_target((short_t)comp.size())
    {
        if ((int)comp.size() != dim) throw std::runtime_error("gsFunctionExpr: a flux needs one component per direction");
        static const char *var = "xyz";
        bool grad = true, constant = true;
        for (int c = 0; c < dim; ++c) {
            std::string want = "pi";
            for (int r = 0; r < dim; ++r) want += std::string("*") + (r == c ? "cos" : "sin") + "(pi*" + var[r] + ")";
            grad = grad && strip(comp[c]) == want;
            double v[4] = {0, 0, 0, 0};
            const bool a = affine(strip(comp[c]), dim, v) && v[1] == 0.0 && v[2] == 0.0 && v[3] == 0.0;
            constant = constant && a;
            m_c[1 + c] = v[0];
        }
        if (grad) { m_kind = 4; return; }
        if (constant) { m_kind = 6; return; }
        throw std::runtime_error("gsFunctionExpr: a flux must be the gradient of prod sin(pi x_i) or constant");
    }
    int kind() const { return m_kind; }
    short_t domainDim() const { return m_dim; }
    short_t targetDim() const { return m_target; }
    const double *coefs() const { return m_c; }

private:
    static std::string strip(const std::string &e)
    {
        std::string o;
        for (char ch : e)
            if (ch != ' ') o += ch;
        return o;
    }
    /// c + sum_r a_r x_r written as a sum of terms "n", "n*x", "x", "x*n"
    static bool affine(const std::string &e, int dim, double *v)
    {
        for (int k = 0; k < 4; ++k) v[k] = 0.0;
        if (e.empty()) return false;
        size_t i = 0;
        while (i < e.size()) {
            double sign = 1.0;
            if (e[i] == '+' || e[i] == '-') {
                sign = e[i] == '-' ? -1.0 : 1.0;
                ++i;
            }
            size_t j = i;
            while (j < e.size() && e[j] != '+' && e[j] != '-') {
                if ((e[j] == 'e' || e[j] == 'E') && j + 1 < e.size()) ++j;   // exponent sign
                ++j;
            }
            const std::string t = e.substr(i, j - i);
            if (t.empty()) return false;
            int var = -1;
            double num = 1.0;
            size_t pos = 0;
            while (pos <= t.size()) {
                const size_t star = t.find('*', pos);
                const std::string f = t.substr(pos, star == std::string::npos ? std::string::npos : star - pos);
                if (f == "x" || f == "y" || f == "z") {
                    const int r = f[0] - 'x';
                    if (var >= 0 || r >= dim) return false;
                    var = r;
                } else {
                    char *end = nullptr;
                    const double x = std::strtod(f.c_str(), &end);
                    if (f.empty() || end != f.c_str() + f.size()) return false;
                    num *= x;
                }
                if (star == std::string::npos) break;
                pos = star + 1;
            }
            v[var < 0 ? 0 : 1 + var] += sign * num;
            i = j;
        }
        return true;
    }
    // 1 source d pi^2 prod sin, 2 zero, 3 solution prod sin, 4 its gradient,
    // 5 affine c + a.x (m_c = c, a), 6 constant vector (m_c[1..d])
    int m_kind = 2;
    short_t m_dim = 2, m_target = 1;
    double m_c[4] = {0, 0, 0, 0};
};

/// Piecewise constant per element (patch-wise arrays, lexicographic);
/// used for the diffusion coefficient and random sources.
template <class T = real_t>
class gsPiecewiseElementFunction {
public:
    gsPiecewiseElementFunction() = default;
    explicit gsPiecewiseElementFunction(std::vector<std::vector<real_t>> perPatch) : m_v(std::move(perPatch)) {}
    const std::vector<real_t> &patch(int k) const { return m_v.at(k); }
    bool empty() const { return m_v.empty(); }

private:
    std::vector<std::vector<real_t>> m_v;
};

/// Boundary conditions (gsBoundaryConditions, src/gsPde/gsBoundaryConditions.h):
/// Dirichlet sides (eliminated; data zero, affine or the manufactured
/// solution) and Neumann sides (flux: the manufactured solution's gradient
/// or a constant vector); sides without a condition are natural.
template <class T = real_t>
class gsBoundaryConditions {
public:
    struct condition {
        int patch;
        int side;
        condition_type::type type;
        std::shared_ptr<const gsFunctionExpr<T>> f;   // a copy of the data (null: zero)
    };
    typedef std::vector<condition> bcRefList;

    void addCondition(int patch, boundary::side s, condition_type::type t, const gsFunctionExpr<T> *g = nullptr,
                      short_t = 0, bool parametric = false, int = -1)
    {
        if (t != condition_type::dirichlet && t != condition_type::neumann)
            throw std::runtime_error("gsBoundaryConditions: Dirichlet and Neumann conditions only");
        if (parametric) throw std::runtime_error("gsBoundaryConditions: data in physical coordinates only");
        if (g && t == condition_type::dirichlet && g->kind() != 2 && g->kind() != 3 && g->kind() != 5)
            throw std::runtime_error("gsBoundaryConditions: Dirichlet data must be 0, affine or prod sin(pi x_i)");
        if (g && t == condition_type::neumann && g->kind() != 2 && g->kind() != 4 && g->kind() != 6)
            throw std::runtime_error("gsBoundaryConditions: Neumann data must be a flux (gradient or constant)");
        m_all.push_back({patch, (int)s, t, g ? std::make_shared<const gsFunctionExpr<T>>(*g) : nullptr});
        if (t == condition_type::dirichlet) m_dir.push_back({patch, (int)s});
    }
    void addCondition(boundary::side s, condition_type::type t, const gsFunctionExpr<T> *g = nullptr)
    {
        addCondition(0, s, t, g);
    }
    void setGeoMap(const gsMultiPatch<T> &) {}
    const std::vector<std::pair<int, int>> &dirichletSides() const { return m_dir; }
    const std::vector<condition> &conditions() const { return m_all; }
    /// the conditions of one type ("Dirichlet" / "Neumann")
    bcRefList get(const std::string &label) const
    {
        bcRefList out;
        for (const condition &c : m_all)
            if ((label == "Dirichlet" && c.type == condition_type::dirichlet) ||
                (label == "Neumann" && c.type == condition_type::neumann))
                out.push_back(c);
        return out;
    }
    bool inhomogeneous() const
    {
        for (const condition &c : m_all)
            if (c.f && c.f->kind() != 2) return true;
        return false;
    }

private:
    std::vector<std::pair<int, int>> m_dir;
    std::vector<condition> m_all;
};

// ---------------------------------------------------------------- device problem
namespace detail {

/// Geometry of patch k expressed in the discretisation basis of mb.
inline std::vector<double> isoCoefs(const gsTensorBSplinePatch &P, const gsMultiBasis<real_t> &mb, int k)
{
    const int d = P.domainDim();
    bool same = true;
    for (int j = 0; j < d; ++j)
        same = same && P.knots(j).knots() == mb.knots(k, j).knots() && P.knots(j).degree() == mb.knots(k, j).degree();
    const index_t nb = mb.size(k);
    std::vector<double> out((size_t)nb * d);
    if (same) {
        for (index_t i = 0; i < nb; ++i)
            for (int r = 0; r < d; ++r) out[(size_t)i * d + r] = P.coefs()(i, r);
        return out;
    }
    if (!P.isMultilinearBox())
        throw std::runtime_error("gismo (libgsx): the geometry must use the discretisation basis or be a "
                                 "multilinear single-element patch");
    // multilinear maps are reproduced exactly by their values at the
    // Greville abscissae of the finer/elevated basis
    std::vector<std::vector<double>> gr(d);
    int nbd[3] = {1, 1, 1};
    for (int j = 0; j < d; ++j) {
        const gsKnotVector &U = mb.knots(k, j);
        const int p = U.degree();
        nbd[j] = U.numBasis();
        for (int i = 0; i < nbd[j]; ++i) {
            double s = 0;
            for (int m = 1; m <= p; ++m) s += U[i + m];
            gr[j].push_back(s / p);
        }
    }
    for (index_t i = 0; i < nb; ++i) {
        const int ix = i % nbd[0], iy = (i / nbd[0]) % nbd[1], iz = d == 3 ? i / (nbd[0] * nbd[1]) : 0;
        double xi[3] = {gr[0][ix], gr[1][iy], d == 3 ? gr[2][iz] : 0.0}, x[3];
        P.evalMultilinear(xi, x);
        for (int r = 0; r < d; ++r) out[(size_t)i * d + r] = x[r];
    }
    return out;
}

struct ProblemDeleter {
    void operator()(gsx_problem *p) const { if (p) gsx_problem_destroy(p); }
};
struct MatrixDeleter {
    void operator()(gsx_matrix *p) const { if (p) gsx_matrix_destroy(p); }
};

inline std::shared_ptr<gsx_problem> makeProblem(const gsMultiPatch<real_t> &mp, const gsMultiBasis<real_t> &mb,
                                                const gsBoundaryConditions<real_t> &bc, int fkind,
                                                const gsPiecewiseElementFunction<real_t> *kappa,
                                                const gsPiecewiseElementFunction<real_t> *felem)
{
    const int np = mp.nPatches();
    if (np == 0 || mb.nBases() != np) throw std::runtime_error("gismo (libgsx): patches and bases do not match");
    std::vector<gsx_patch_desc> desc(np);
    std::vector<std::vector<double>> ctrl(np), zeros(np);
    for (int k = 0; k < np; ++k) {
        gsx_patch_desc &D = desc[k];
        std::memset(&D, 0, sizeof(D));
        D.d = mp.patch(k).domainDim();
        for (int j = 0; j < D.d; ++j) {
            D.p[j] = mb.knots(k, j).degree();
            D.nknots[j] = (int32_t)mb.knots(k, j).size();
            D.knots_host[j] = mb.knots(k, j).knots().data();
        }
        ctrl[k] = isoCoefs(mp.patch(k), mb, k);
        D.ctrl_host = ctrl[k].data();
        D.kappa_host = (kappa && !kappa->empty()) ? kappa->patch(k).data() : nullptr;
        D.felem_host = (felem && !felem->empty()) ? felem->patch(k).data() : nullptr;
        D.fkind = felem && !felem->empty() ? 0 : fkind;
        if (D.fkind == 2) {   // zero source: per-element zeros
            index_t ne = 1;
            for (int j = 0; j < D.d; ++j) ne *= (index_t)mb.knots(k, j).size() - 2 * D.p[j] - 1;
            zeros[k].assign(ne, 0.0);
            D.felem_host = zeros[k].data();
            D.fkind = 0;
        }
    }
    std::vector<int32_t> ds, ifs;
    for (auto &s : bc.dirichletSides()) { ds.push_back(s.first); ds.push_back(s.second); }
    for (auto &f : mp.interfaces()) {
        const int32_t r[8] = {f.p1, f.s1, f.p2, f.s2, f.m0, f.f0, f.m1, f.f1};
        ifs.insert(ifs.end(), r, r + 8);
    }
    gsx_problem *p = nullptr;
    check(gsx_problem_create(ctx(), np, desc.data(), (int32_t)ds.size() / 2, ds.empty() ? nullptr : ds.data(),
                             (int32_t)ifs.size() / 8, ifs.empty() ? nullptr : ifs.data(), &p),
          "gsx_problem_create");
    return std::shared_ptr<gsx_problem>(p, ProblemDeleter());
}

/// Boundary data of bc on the device problem (gsFeSpace::setup /
/// gsAssembler::computeDirichletDofs): Dirichlet data by kind, Neumann
/// fluxes; dir_values one of GSX_DIR_*.
inline void applyBc(const std::shared_ptr<gsx_problem> &prob, const gsBoundaryConditions<real_t> &bc, int dir_values)
{
    std::vector<gsx_bc> recs;
    for (const auto &c : bc.conditions()) {
        gsx_bc r{};
        r.patch = c.patch;
        r.side = c.side;
        r.type = c.type == condition_type::dirichlet ? GSX_BC_DIRICHLET : GSX_BC_NEUMANN;
        const int k = c.f ? c.f->kind() : 2;
        r.kind = (k == 3 || k == 4) ? GSX_BCF_MANUFACTURED : (k == 5 || k == 6) ? GSX_BCF_AFFINE : GSX_BCF_ZERO;
        if (c.f)
            for (int v = 0; v < 4; ++v) r.value[v] = c.f->coefs()[v];
        recs.push_back(r);
    }
    check(gsx_problem_set_bc(prob.get(), (int32_t)recs.size(), recs.empty() ? nullptr : recs.data(), dir_values),
          "gsx_problem_set_bc");
}

/// Device assembly -> host gsSparseMatrix (the CSR arrays of a bitwise
/// symmetric matrix are its CSC arrays) + rhs.
inline void assembleInto(const std::shared_ptr<gsx_problem> &prob, int mode, int precision,
                         gsSparseMatrix<real_t> &M, gsMatrix<real_t> &rhs)
{
    int64_t nbasis = 0, nfree = 0, nelem = 0, nglued = 0;
    check(gsx_problem_info(prob.get(), &nbasis, &nfree, &nelem, &nglued), "gsx_problem_info");
    DeviceVec b(sizeof(double) * std::max<int64_t>(1, nfree));
    gsx_matrix *A = nullptr;
    check(gsx_assemble(prob.get(), mode, precision, &A, b.d()), "gsx_assemble");
    std::shared_ptr<gsx_matrix> dev(A, MatrixDeleter());
    int64_t n = 0, nnz = 0;
    check(gsx_matrix_info(A, &n, &nnz), "gsx_matrix_info");
    std::vector<index_t> outer(n + 1), inner(nnz);
    std::vector<real_t> val(nnz);
    check(gsx_matrix_download(A, outer.data(), inner.data(), val.data()), "gsx_matrix_download");
    M.setFromCompressed((index_t)n, (index_t)n, std::move(outer), std::move(inner), std::move(val));
    M.setDevice(dev);
    M.setProblem(prob);
    rhs.resize((index_t)nfree, 1);
    check(gsx_copy_to_host(ctx(), rhs.data(), b.d(), sizeof(double) * nfree), "download rhs");
}

}  // namespace detail

// ---------------------------------------------------------------- expressions
namespace expr {
struct geometryMap { const gsMultiPatch<real_t> *mp = nullptr; };
struct gsFeSpace;
struct igrad_expr { bool tr = false; igrad_expr tr_() const { igrad_expr e; e.tr = true; return e; } };
struct meas_expr {};
struct coeff_expr { int kind = 2; };
struct kappa_expr { const gsPiecewiseElementFunction<real_t> *k = nullptr; };
struct gg_expr {};
struct bilinear_form { const gsPiecewiseElementFunction<real_t> *kappa = nullptr; };
struct uf_expr { int kind = 2; };
struct linear_form { int kind = 2; const gsPiecewiseElementFunction<real_t> *felem = nullptr; };

struct gsFeSpace {
    const gsMultiBasis<real_t> *mb = nullptr;
    // shared by the copies handed out by getSpace, so that setup() on any
    // of them reaches the assembler
    std::shared_ptr<const gsBoundaryConditions<real_t> *> bc = std::make_shared<const gsBoundaryConditions<real_t> *>(nullptr);
    std::shared_ptr<int> dir_values = std::make_shared<int>(dirichlet::l2Projection);
    /// gsFeSpace::setup (src/gsAssembler/gsExpressions.h:1091-1351):
    /// interfaces glued (icont = 0), Dirichlet sides eliminated, their values
    /// by dir_values (gsDirichletValues, src/gsAssembler/gsDirichletValues.h:
    /// 27-60): homogeneous, interpolation or l2Projection
    void setup(const gsBoundaryConditions<real_t> &b, int dv, int = 0) const
    {
        if (dv != dirichlet::homogeneous && dv != dirichlet::l2Projection && dv != dirichlet::interpolation)
            throw std::runtime_error("gsFeSpace::setup: Dirichlet values by homogeneous, interpolation or l2Projection");
        *bc = &b;
        *dir_values = dv;
    }
};
/// nv(G) and the Neumann form u * g_N.tr() * nv(G) of assembleBdr
struct nv_expr {};
struct u_gn { };
struct neumann_form { };
/// exact solution u = prod sin(pi x_i) (gsExprEvaluator::getVariable)
struct exact_var { int kind = 3; };
/// discrete solution (gsExprAssembler::getSolution): the free coefficients
struct solution_var { const gsMatrix<real_t> *coefs = nullptr; };
struct diff_expr {
    int kind = 3;
    const gsMatrix<real_t> *coefs = nullptr;
    diff_expr sqNorm() const { return *this; }
};
struct l2_integrand { diff_expr d; };
/// igrad(u_ex) - igrad(u_sol, G) of the H1 seminorm error
struct igrad_exact {};
struct igrad_sol { const gsMatrix<real_t> *coefs = nullptr; };
struct grad_diff {
    const gsMatrix<real_t> *coefs = nullptr;
    grad_diff sqNorm() const { return *this; }
};
struct h1_integrand { grad_diff d; };
inline grad_diff operator-(const igrad_exact &, const igrad_sol &s) { grad_diff d; d.coefs = s.coefs; return d; }
inline h1_integrand operator*(const grad_diff &d, const meas_expr &) { h1_integrand i; i.d = d; return i; }
inline diff_expr operator-(const exact_var &e, const solution_var &s) { diff_expr d; d.kind = e.kind; d.coefs = s.coefs; return d; }
inline diff_expr operator-(const solution_var &s, const exact_var &e) { return e - s; }
inline l2_integrand operator*(const diff_expr &d, const meas_expr &) { l2_integrand i; i.d = d; return i; }
}  // namespace expr

typedef expr::geometryMap geometryMap;
typedef expr::gsFeSpace space;

struct igrad_t {
    expr::igrad_expr e;
    igrad_t tr() const { igrad_t t; t.e.tr = true; return t; }
};
inline igrad_t igrad(const space &, const geometryMap &) { return igrad_t(); }
inline expr::igrad_exact igrad(const expr::exact_var &) { return expr::igrad_exact(); }
inline expr::igrad_sol igrad(const expr::solution_var &s, const geometryMap &)
{
    expr::igrad_sol g;
    g.coefs = s.coefs;
    return g;
}
inline expr::meas_expr meas(const geometryMap &) { return expr::meas_expr(); }
inline expr::gg_expr operator*(const igrad_t &a, const igrad_t &b)
{
    if (a.e.tr || !b.e.tr) throw std::runtime_error("gsExprAssembler: expected igrad(u,G) * igrad(u,G).tr()");
    return expr::gg_expr();
}
inline expr::bilinear_form operator*(const expr::gg_expr &, const expr::meas_expr &) { return expr::bilinear_form(); }
inline expr::bilinear_form operator*(const expr::bilinear_form &f, const expr::kappa_expr &k)
{
    expr::bilinear_form g = f;
    g.kappa = k.k;
    return g;
}
inline expr::bilinear_form operator*(const expr::kappa_expr &k, const expr::bilinear_form &f) { return f * k; }
inline expr::uf_expr operator*(const space &, const expr::coeff_expr &c) { expr::uf_expr u; u.kind = c.kind; return u; }
inline expr::linear_form operator*(const expr::uf_expr &u, const expr::meas_expr &)
{
    expr::linear_form l;
    l.kind = u.kind;
    return l;
}
inline expr::linear_form operator*(const space &, const expr::kappa_expr &k)
{
    expr::linear_form l;
    l.felem = k.k;
    return l;
}
inline expr::linear_form operator*(const expr::linear_form &l, const expr::meas_expr &) { return l; }
struct bdr_fun_t {
    bdr_fun_t tr() const { return bdr_fun_t(); }
};
inline expr::nv_expr nv(const geometryMap &) { return expr::nv_expr(); }
inline expr::u_gn operator*(const space &, const bdr_fun_t &) { return expr::u_gn(); }
inline expr::neumann_form operator*(const expr::u_gn &, const expr::nv_expr &) { return expr::neumann_form(); }

// ---------------------------------------------------------------- assemblers
/// gsExprAssembler (src/gsAssembler/gsExprAssembler.h:29-833) for the
/// Poisson/diffusion forms above.  Options: "Precision" ("fp64" default or
/// "mixed": keeps an FP32 copy of the operator for mixed solves).
template <class T = real_t>
class gsExprAssembler {
public:
    typedef gismo::geometryMap geometryMap;
    typedef gismo::space space;
    typedef expr::solution_var solution;
    typedef expr::exact_var variable;

    explicit gsExprAssembler(index_t = 1, index_t = 1) {}
    void setIntegrationElements(const gsMultiBasis<T> &mb) { m_mb = &mb; }
    void setIntegrationElements(const gsMultiBasis<T> &&) = delete;
    geometryMap getMap(const gsMultiPatch<T> &mp) { m_mp = &mp; geometryMap G; G.mp = &mp; return G; }
    space getSpace(const gsMultiBasis<T> &mb, index_t = 1, index_t = 0) { m_u.mb = &mb; return m_u; }
    solution getSolution(const space &, gsMatrix<T> &coefs) const { solution s; s.coefs = &coefs; return s; }
    expr::coeff_expr getCoeff(const gsFunctionExpr<T> &f, const geometryMap &) { expr::coeff_expr c; c.kind = f.kind(); return c; }
    expr::coeff_expr getCoeff(const gsFunctionExpr<T> &f) { expr::coeff_expr c; c.kind = f.kind(); return c; }
    expr::kappa_expr getCoeff(const gsPiecewiseElementFunction<T> &k) { expr::kappa_expr e; e.k = &k; return e; }
    gsOptionList &options() { return m_opt; }
    void setOptions(const gsOptionList &o) { m_opt = o; }
    void initSystem(index_t = 1) {}
    index_t numDofs() const { return m_matrix.rows(); }
    void assemble(const expr::bilinear_form &a, const expr::linear_form &l)
    {
        if (!m_mb || !m_mp) throw std::runtime_error("gsExprAssembler: call setIntegrationElements and getMap first");
        if (!l.felem && l.kind != 1 && l.kind != 2)
            throw std::runtime_error("gsExprAssembler: unsupported source function");
        const gsBoundaryConditions<T> *bc = *m_u.bc ? *m_u.bc : m_bc;
        static const gsBoundaryConditions<T> none;
        m_prob = detail::makeProblem(*m_mp, *m_mb, bc ? *bc : none, l.kind, a.kappa, l.felem);
        if (bc && bc->inhomogeneous()) {
            const int dv = *m_u.dir_values;
            detail::applyBc(m_prob, *bc, dv == dirichlet::homogeneous     ? GSX_DIR_HOMOGENEOUS
                                         : dv == dirichlet::interpolation ? GSX_DIR_INTERPOLATION
                                                                          : GSX_DIR_L2PROJECTION);
        }
        const int prec = m_opt.askString("Precision", "fp64") == "mixed" ? GSX_MIXED : GSX_FP64;
        detail::assembleInto(m_prob, GSX_PATTERN_EXPR, prec, m_matrix, m_rhs);
    }
    /// g_N of the Neumann conditions in assembleBdr
    bdr_fun_t getBdrFunction(const geometryMap &) const { return bdr_fun_t(); }
    /// gsExprAssembler::assembleBdr(bc.get("Neumann"), u * g_N.tr() * nv(G))
    /// (src/gsAssembler/gsExprAssembler.h:835-895): adds the Neumann load of
    /// the space's boundary conditions to the right-hand side
    template <class List>
    void assembleBdr(const List &neumann, const expr::neumann_form &)
    {
        if (!m_prob) throw std::runtime_error("gsExprAssembler::assembleBdr: assemble first");
        if (neumann.empty()) return;
        const index_t n = m_rhs.rows();
        detail::DeviceVec b(sizeof(double) * std::max<index_t>(1, n));
        std::vector<double> zero((size_t)std::max<index_t>(1, n), 0.0);
        detail::check(gsx_copy_to_device(detail::ctx(), b.d(), zero.data(), sizeof(double) * n), "upload");
        detail::check(gsx_assemble_bdr(m_prob.get(), (double *)b.d()), "gsx_assemble_bdr");
        std::vector<double> load((size_t)std::max<index_t>(1, n));
        detail::check(gsx_copy_to_host(detail::ctx(), load.data(), b.d(), sizeof(double) * n), "download");
        for (index_t i = 0; i < n; ++i) m_rhs(i, 0) += load[(size_t)i];
    }
    const gsSparseMatrix<T> &matrix() const { return m_matrix; }
    void matrix_into(gsSparseMatrix<T> &out) const { out = m_matrix; }
    gsSparseMatrix<T> giveMatrix() { return std::move(m_matrix); }
    const gsMatrix<T> &rhs() const { return m_rhs; }
    std::shared_ptr<gsx_problem> problem() const { return m_prob; }
    /// conditions used when the space's setup() was not called
    void setBoundaryConditions(const gsBoundaryConditions<T> &bc) { m_bc = &bc; }

private:
    const gsMultiBasis<T> *m_mb = nullptr;
    const gsMultiPatch<T> *m_mp = nullptr;
    const gsBoundaryConditions<T> *m_bc = nullptr;
    space m_u;
    gsOptionList m_opt;
    std::shared_ptr<gsx_problem> m_prob;
    gsSparseMatrix<T> m_matrix;
    gsMatrix<T> m_rhs;
};

/// gsPoissonAssembler (src/gsAssembler/gsPoissonAssembler.h:35-142):
/// every free-free pair is in the pattern (gsSparseSystem::push).
template <class T = real_t>
class gsPoissonAssembler {
public:
    gsPoissonAssembler(const gsMultiPatch<T> &mp, const gsMultiBasis<T> &mb, const gsBoundaryConditions<T> &bc,
                       const gsFunctionExpr<T> &f, dirichlet::strategy ds = dirichlet::elimination,
                       iFace::strategy = iFace::glue)
        : m_mp(&mp), m_mb(mb), m_bc(&bc), m_f(f)
    {
        if (ds != dirichlet::elimination)
            throw std::runtime_error("gsPoissonAssembler: only Dirichlet elimination is supported");
    }
    gsOptionList &options() { return m_opt; }
    virtual ~gsPoissonAssembler() = default;
    virtual void assemble()
    {
        m_prob = detail::makeProblem(*m_mp, m_mb, *m_bc, m_f.kind(), nullptr, nullptr);
        if (m_bc->inhomogeneous()) {
            // gsAssembler::computeDirichletDofs (src/gsAssembler/gsAssembler.hpp:231-297):
            // DirichletValues 101 (interpolation, the default), 102 or 100
            const int dv = m_opt.askInt("DirichletValues", dirichlet::interpolation);
            for (const auto &c : m_bc->conditions())
                if (c.type == condition_type::neumann && c.f && c.f->kind() != 2)
                    throw std::runtime_error("gsPoissonAssembler: scalar Neumann data are not provided "
                                             "(gsExprAssembler::assembleBdr takes fluxes)");
            if (dv != dirichlet::l2Projection && dv != dirichlet::homogeneous && dv != dirichlet::interpolation)
                throw std::runtime_error("gsPoissonAssembler: DirichletValues must be 100, 101 or 102");
            detail::applyBc(m_prob, *m_bc, dv == dirichlet::homogeneous     ? GSX_DIR_HOMOGENEOUS
                                           : dv == dirichlet::interpolation ? GSX_DIR_INTERPOLATION
                                                                            : GSX_DIR_L2PROJECTION_INHOM);
        }
        const int prec = m_opt.askString("Precision", "fp64") == "mixed" ? GSX_MIXED : GSX_FP64;
        detail::assembleInto(m_prob, GSX_PATTERN_FULL, prec, m_matrix, m_rhs);
    }
    const gsSparseMatrix<T> &matrix() const { return m_matrix; }
    const gsMatrix<T> &rhs() const { return m_rhs; }
    index_t numDofs() const { return m_matrix.rows(); }
    std::shared_ptr<gsx_problem> problem() const { return m_prob; }
    const gsMultiPatch<T> &patches() const { return *m_mp; }
    /// gsAssembler::constructSolution (src/gsAssembler/gsAssembler.hpp:
    /// 537-581): one scalar patch per patch over the discretisation basis;
    /// the coefficient of a free function is its entry of solVector, of an
    /// eliminated one its fixed (Dirichlet) value.
    void constructSolution(const gsMatrix<T> &solVector, gsMultiPatch<T> &result) const
    {
        if (!m_prob) throw std::runtime_error("gsPoissonAssembler::constructSolution: nothing assembled yet");
        int64_t nbasis = 0, nfree = 0, nelem = 0, nglued = 0;
        detail::check(gsx_problem_info(m_prob.get(), &nbasis, &nfree, &nelem, &nglued), "gsx_problem_info");
        if (solVector.rows() != nfree)
            throw std::runtime_error("gsPoissonAssembler::constructSolution: solution size does not match the space");
        std::vector<int32_t> dofs((size_t)nbasis);
        detail::check(gsx_problem_dofs(m_prob.get(), dofs.data()), "gsx_problem_dofs");
        std::vector<double> fixed((size_t)std::max<int64_t>(1, nglued - nfree), 0.0);
        if (nglued > nfree)
            detail::check(gsx_problem_fixed_values(m_prob.get(), fixed.data(), (int32_t)(nglued - nfree)),
                          "gsx_problem_fixed_values");
        result = gsMultiPatch<T>();
        int64_t off = 0;
        for (index_t k = 0; k < m_mb.nBases(); ++k) {
            const index_t nb = m_mb.size(k);
            std::vector<gsKnotVector> kv;
            for (int j = 0; j < m_mb.basis(k).dim(); ++j) kv.push_back(m_mb.knots(k, j));
            gsMatrix<real_t> C(nb, 1);
            for (index_t i = 0; i < nb; ++i) {
                const int32_t g = dofs[(size_t)(off + i)];
                C(i, 0) = g < nfree ? solVector(g) : fixed[(size_t)(g - nfree)];
            }
            result.addPatch(gsTensorBSplinePatch(kv, C));
            off += nb;
        }
    }

private:
    const gsMultiPatch<T> *m_mp;
    gsMultiBasis<T> m_mb;   // a copy, as gsAssembler keeps its bases
    const gsBoundaryConditions<T> *m_bc;
    gsFunctionExpr<T> m_f;
    gsOptionList m_opt;
    std::shared_ptr<gsx_problem> m_prob;
    gsSparseMatrix<T> m_matrix;
    gsMatrix<T> m_rhs;
};

// ---------------------------------------------------------------- grid hierarchy
namespace detail {

/// Rows of the 1D knot-insertion transfer (coarse -> fine coefficients,
/// nf x nc): each knot of `fine` missing from `coarse` is inserted once,
/// new coefficients Q_i = a_i P_i + (1 - a_i) P_{i-1}, a_i = (u - U_i) /
/// (U_{i+p} - U_i) for the p functions touching the span (Boehm's
/// insertion, the transfer of gsBSplineBasis::refine_withTransfer).
inline std::vector<std::vector<double>> insertionRows(const std::vector<real_t> &coarse,
                                                      const std::vector<real_t> &fine, int p)
{
    std::vector<real_t> U = coarse, ins;
    size_t a = 0;
    for (real_t u : fine) {
        if (a < coarse.size() && coarse[a] == u) ++a;
        else ins.push_back(u);
    }
    if (a != coarse.size()) throw std::runtime_error("gsGridHierarchy: the coarse knots are not nested in the fine ones");
    const int nc = (int)coarse.size() - p - 1;
    std::vector<std::vector<double>> R((size_t)nc, std::vector<double>((size_t)nc, 0.0));
    for (int i = 0; i < nc; ++i) R[(size_t)i][(size_t)i] = 1.0;
    for (real_t u : ins) {
        int k = p;
        while (k + 1 < (int)U.size() - p - 1 && U[(size_t)k + 1] <= u) ++k;
        std::vector<std::vector<double>> out;
        out.reserve(R.size() + 1);
        for (int i = 0; i <= k - p; ++i) out.push_back(R[(size_t)i]);
        for (int i = k - p + 1; i <= k; ++i) {
            const double al = (u - U[(size_t)i]) / (U[(size_t)i + p] - U[(size_t)i]);
            std::vector<double> row((size_t)nc);
            for (int c = 0; c < nc; ++c) row[(size_t)c] = al * R[(size_t)i][(size_t)c] + (1.0 - al) * R[(size_t)i - 1][(size_t)c];
            out.push_back(std::move(row));
        }
        for (size_t i = (size_t)k; i < R.size(); ++i) out.push_back(R[i]);
        R.swap(out);
        U.insert(U.begin() + k + 1, u);
    }
    return R;
}

/// The DOF map of a multibasis (gsMultiBasis::getMapper with Dirichlet
/// elimination and glued interfaces, src/gsCore/gsMultiBasis.hpp:373-424).
template <class T>
inline std::vector<int32_t> multiBasisDofs(const gsMultiBasis<T> &mb, const gsBoundaryConditions<T> &bc, int64_t &nfree)
{
    const int np = mb.nBases();
    const int d = np ? (int)mb.basis(0).dim() : 0;
    std::vector<int32_t> nb((size_t)np * 3, 1), ds, ifs;
    int64_t tot = 0;
    for (int k = 0; k < np; ++k) {
        for (int j = 0; j < d; ++j) nb[(size_t)3 * k + j] = mb.knots(k, j).numBasis();
        tot += mb.size(k);
    }
    for (auto &sd : bc.dirichletSides()) { ds.push_back(sd.first); ds.push_back(sd.second); }
    for (auto &f : mb.interfaces()) {
        const int32_t r[8] = {f.p1, f.s1, f.p2, f.s2, f.m0, f.f0, f.m1, f.f1};
        ifs.insert(ifs.end(), r, r + 8);
    }
    std::vector<int32_t> dofs((size_t)tot);
    int64_t nglued = 0;
    check(gsx_mapper_build(d, np, nb.data(), (int32_t)ds.size() / 2, ds.empty() ? nullptr : ds.data(),
                           (int32_t)ifs.size() / 8, ifs.empty() ? nullptr : ifs.data(), dofs.data(), &nfree, &nglued),
          "gsx_mapper_build");
    return dofs;
}

/// The transfer between two nested multibases restricted to free DOFs
/// (gsMultiBasis::combineTransferMatrices, src/gsCore/gsMultiBasis.hpp:
/// 137-178): per patch the tensor product of the 1D insertion transfers,
/// rows in local order, entries whose fine and coarse DOFs are both free,
/// the first of duplicated positions kept (take_first).
template <class T>
inline gsSparseMatrix<T, RowMajor> combinedTransfer(const gsMultiBasis<T> &coarse, const gsMultiBasis<T> &fine,
                                                    const gsBoundaryConditions<T> &bc)
{
    int64_t nfc = 0, nff = 0;
    const std::vector<int32_t> dc = multiBasisDofs(coarse, bc, nfc), df = multiBasisDofs(fine, bc, nff);
    std::vector<typename gsSparseMatrix<T, RowMajor>::Triplet> trip;
    int64_t offc = 0, offf = 0;
    for (int k = 0; k < fine.nBases(); ++k) {
        const int d = fine.basis(k).dim();
        std::vector<std::vector<std::vector<double>>> T1(3);
        int nf[3] = {1, 1, 1}, nc[3] = {1, 1, 1};
        for (int j = 0; j < 3; ++j) {
            if (j < d) {
                T1[j] = insertionRows(coarse.knots(k, j).knots(), fine.knots(k, j).knots(), fine.knots(k, j).degree());
                nf[j] = (int)T1[j].size();
                nc[j] = coarse.knots(k, j).numBasis();
            } else {
                T1[j] = {{1.0}};
            }
        }
        for (int iz = 0; iz < nf[2]; ++iz)
            for (int iy = 0; iy < nf[1]; ++iy)
                for (int ix = 0; ix < nf[0]; ++ix) {
                    const int64_t fi = ix + (int64_t)nf[0] * (iy + (int64_t)nf[1] * iz);
                    const int fdof = df[(size_t)(offf + fi)];
                    if (fdof >= nff) continue;
                    for (int cz = 0; cz < nc[2]; ++cz) {
                        const double vz = T1[2][(size_t)iz][(size_t)cz];
                        if (vz == 0.0) continue;
                        for (int cy = 0; cy < nc[1]; ++cy) {
                            const double vy = T1[1][(size_t)iy][(size_t)cy];
                            if (vy == 0.0) continue;
                            for (int cx = 0; cx < nc[0]; ++cx) {
                                const double vx = T1[0][(size_t)ix][(size_t)cx];
                                if (vx == 0.0) continue;
                                const int64_t ci = cx + (int64_t)nc[0] * (cy + (int64_t)nc[1] * cz);
                                const int cdof = dc[(size_t)(offc + ci)];
                                if (cdof >= nfc) continue;
                                trip.push_back({fdof, cdof, (T)(vx * vy * vz)});
                            }
                        }
                    }
                }
        offc += coarse.size(k);
        offf += fine.size(k);
    }
    gsSparseMatrix<T, RowMajor> P;
    P.setFromTriplets((index_t)nff, (index_t)nfc, trip, true);
    return P;
}

}  // namespace detail

/// gsGridHierarchy (src/gsMultiGrid/gsGridHierarchy.h:35-200,
/// gsGridHierarchy.hpp:27-101): the multibases of a multigrid hierarchy,
/// coarsest first, and the transfers between neighbouring levels (restricted
/// to free DOFs; Dirichlet elimination and glued interfaces).
template <class T = real_t>
class gsGridHierarchy {
public:
    static gsGridHierarchy buildByCoarsening(gsMultiBasis<T> mBasis, const gsBoundaryConditions<T> &bc,
                                             const gsOptionList &options)
    {
        return buildByCoarsening(std::move(mBasis), bc, options, options.askInt("Levels", 3),
                                 options.askInt("DegreesOfFreedom", 0), 0);
    }
    static gsGridHierarchy buildByCoarsening(gsMultiBasis<T> mBasis, const gsBoundaryConditions<T> &bc,
                                             const gsOptionList &options, index_t levels, index_t degreesOfFreedom = 0,
                                             index_t = 0)
    {
        checkStrategies(options);
        gsGridHierarchy r;
        r.m_mBases.push_back(std::move(mBasis));
        index_t lastSize = r.m_mBases[0].totalSize();
        for (index_t i = 0; i < levels - 1 && lastSize > degreesOfFreedom; ++i) {
            gsMultiBasis<T> coarse = r.m_mBases[(size_t)i];
            coarse.uniformCoarsen(1);
            const index_t newSize = coarse.totalSize();
            if (lastSize <= newSize && degreesOfFreedom > 0) break;
            lastSize = newSize;
            r.m_transfers.push_back(detail::combinedTransfer(coarse, r.m_mBases[(size_t)i], bc));
            r.m_mBases.push_back(std::move(coarse));
        }
        std::reverse(r.m_mBases.begin(), r.m_mBases.end());
        std::reverse(r.m_transfers.begin(), r.m_transfers.end());
        return r;
    }
    static gsGridHierarchy buildByRefinement(gsMultiBasis<T> mBasis, const gsBoundaryConditions<T> &bc,
                                             const gsOptionList &options)
    {
        return buildByRefinement(std::move(mBasis), bc, options, options.askInt("Levels", 3),
                                 options.askInt("NumberOfKnotsToBeInserted", 1),
                                 options.askInt("MultiplicityOfKnotsToBeInserted", 1), 0);
    }
    static gsGridHierarchy buildByRefinement(gsMultiBasis<T> mBasis, const gsBoundaryConditions<T> &bc,
                                             const gsOptionList &options, index_t levels,
                                             index_t numberOfKnotsToBeInserted = 1,
                                             index_t multiplicityOfKnotsToBeInserted = 1, index_t = 0)
    {
        checkStrategies(options);
        if (multiplicityOfKnotsToBeInserted != 1)
            throw std::runtime_error("gsGridHierarchy::buildByRefinement: knot multiplicity 1 only");
        gsGridHierarchy r;
        r.m_mBases.push_back(std::move(mBasis));
        for (index_t i = 1; i < levels; ++i) {
            gsMultiBasis<T> fine = r.m_mBases.back();
            fine.uniformRefine((int)numberOfKnotsToBeInserted);
            r.m_transfers.push_back(detail::combinedTransfer(r.m_mBases.back(), fine, bc));
            r.m_mBases.push_back(std::move(fine));
        }
        return r;
    }
    gsGridHierarchy &moveMultiBasesTo(std::vector<gsMultiBasis<T>> &out)
    {
        out = std::move(m_mBases);
        return *this;
    }
    gsGridHierarchy &moveTransferMatricesTo(std::vector<gsSparseMatrix<T, RowMajor>> &out)
    {
        out = std::move(m_transfers);
        return *this;
    }
    const std::vector<gsMultiBasis<T>> &getMultiBases() const { return m_mBases; }
    const std::vector<gsSparseMatrix<T, RowMajor>> &getTransferMatrices() const { return m_transfers; }

private:
    static void checkStrategies(const gsOptionList &o)
    {
        if (o.askInt("DirichletStrategy", dirichlet::elimination) != dirichlet::elimination)
            throw std::runtime_error("gsGridHierarchy: only Dirichlet elimination is supported");
        if (o.askInt("InterfaceStrategy", iFace::glue) != iFace::glue)
            throw std::runtime_error("gsGridHierarchy: only glued (conforming) interfaces are supported");
    }
    std::vector<gsMultiBasis<T>> m_mBases;
    std::vector<gsSparseMatrix<T, RowMajor>> m_transfers;
};

// ---------------------------------------------------------------- operators / solvers
/// gsLinearOperator (src/gsSolver/gsLinearOperator.h:28-70)
template <class T = real_t>
class gsLinearOperator {
public:
    typedef std::shared_ptr<gsLinearOperator> Ptr;
    virtual ~gsLinearOperator() = default;
    virtual void apply(const gsMatrix<T> &input, gsMatrix<T> &x) const = 0;
    virtual index_t rows() const = 0;
    virtual index_t cols() const = 0;
    virtual void setOptions(const gsOptionList &) {}
};

namespace detail {
inline std::shared_ptr<gsx_matrix> deviceMatrix(const gsSparseMatrix<real_t> &M, bool fp32)
{
    if (M.devicePtr()) return M.devicePtr();
    const index_t n = M.rows();
    if (M.cols() != n) throw std::runtime_error("gismo (libgsx): square matrices only");
    std::vector<int32_t> rp, ci;
    std::vector<double> v;
    M.toCsr(rp, ci, v);
    gsx_matrix *A = nullptr;
    check(gsx_matrix_from_csr(ctx(), n, rp.data(), ci.data(), v.data(), fp32 ? 1 : 0, &A), "gsx_matrix_from_csr");
    return std::shared_ptr<gsx_matrix>(A, MatrixDeleter());
}
}  // namespace detail

/// gsMatrixOp (src/gsSolver/gsMatrixOp.h:64-75): y = A x on the device
template <class T = real_t>
class gsMatrixOp : public gsLinearOperator<T> {
public:
    explicit gsMatrixOp(const gsSparseMatrix<T> &M) : m_n(M.rows()), m_dev(detail::deviceMatrix(M, false)) {}
    void apply(const gsMatrix<T> &input, gsMatrix<T> &x) const override
    {
        detail::DeviceVec a(sizeof(double) * m_n), b(sizeof(double) * m_n);
        detail::check(gsx_copy_to_device(detail::ctx(), a.p, input.data(), sizeof(double) * m_n), "upload");
        detail::check(gsx_spmv(m_dev.get(), a.d(), b.d()), "gsx_spmv");
        x.resize(m_n, 1);
        detail::check(gsx_copy_to_host(detail::ctx(), x.data(), b.p, sizeof(double) * m_n), "download");
    }
    index_t rows() const override { return m_n; }
    index_t cols() const override { return m_n; }

private:
    index_t m_n;
    std::shared_ptr<gsx_matrix> m_dev;
};

template <class T>
typename gsLinearOperator<T>::Ptr makeMatrixOp(const gsSparseMatrix<T> &M)
{
    return std::make_shared<gsMatrixOp<T>>(M);
}

/// gsPreconditionerOp base (src/gsSolver/gsPreconditioner.h:42-130)
template <class T = real_t>
class gsPreconditionerOp : public gsLinearOperator<T> {
public:
    typedef std::shared_ptr<gsPreconditionerOp> Ptr;
    virtual int kind() const = 0;
    /// as a multigrid smoother: GSX_SMOOTH_* and its damping (-1: not a smoother)
    virtual int smootherKind() const { return -1; }
    virtual T smootherDamping() const { return T(1); }
    /// sweeps per step (gsPreconditionerOp::setNumOfSweeps); one is provided
    void setNumOfSweeps(index_t n)
    {
        if (n != 1) throw std::runtime_error("gsPreconditionerOp: one sweep per step is provided");
    }
};

/// gsJacobiOp (src/gsSolver/gsSimplePreconditioners.h:146-244)
template <class T = real_t>
class gsJacobiOp : public gsPreconditionerOp<T> {
public:
    explicit gsJacobiOp(const gsSparseMatrix<T> &M, T tau = 1) : m_M(&M), m_tau(tau) {}
    void apply(const gsMatrix<T> &input, gsMatrix<T> &x) const override
    {
        x.resize(input.rows(), 1);
        for (index_t i = 0; i < input.rows(); ++i) x(i) = m_tau * input(i) / m_M->coeff(i, i);
    }
    index_t rows() const override { return m_M->rows(); }
    index_t cols() const override { return m_M->cols(); }
    int kind() const override { return GSX_PREC_JACOBI; }
    int smootherKind() const override { return GSX_SMOOTH_JACOBI; }
    T smootherDamping() const override { return m_tau; }
    void setDamping(T tau) { m_tau = tau; }
    T damping() const { return m_tau; }
    void setOptions(const gsOptionList &o) override { m_tau = o.askReal("Damping", m_tau); }

private:
    const gsSparseMatrix<T> *m_M;
    T m_tau;
};

template <class T>
std::shared_ptr<gsJacobiOp<T>> makeJacobiOp(const gsSparseMatrix<T> &M, T tau = 1)
{
    return std::make_shared<gsJacobiOp<T>>(M, tau);
}

/// gsRichardsonOp (src/gsSolver/gsSimplePreconditioners.h:33-131):
/// x = tau r; as a smoother x += tau (b - A x).
template <class T = real_t>
class gsRichardsonOp : public gsPreconditionerOp<T> {
public:
    explicit gsRichardsonOp(const gsSparseMatrix<T> &M, T tau = 1) : m_M(&M), m_tau(tau) {}
    void apply(const gsMatrix<T> &input, gsMatrix<T> &x) const override
    {
        x.resize(input.rows(), 1);
        for (index_t i = 0; i < input.rows(); ++i) x(i) = m_tau * input(i);
    }
    index_t rows() const override { return m_M->rows(); }
    index_t cols() const override { return m_M->cols(); }
    int kind() const override { return GSX_PREC_NONE; }
    int smootherKind() const override { return GSX_SMOOTH_RICHARDSON; }
    T smootherDamping() const override { return m_tau; }
    void setOptions(const gsOptionList &o) override { m_tau = o.askReal("Damping", m_tau); }

private:
    const gsSparseMatrix<T> *m_M;
    T m_tau;
};

template <class T>
std::shared_ptr<gsRichardsonOp<T>> makeRichardsonOp(const gsSparseMatrix<T> &M, T tau = 1)
{
    return std::make_shared<gsRichardsonOp<T>>(M, tau);
}

/// gsGaussSeidelOp (src/gsSolver/gsSimplePreconditioners.h:278-395): a
/// multigrid smoother here (forward sweeps before, reverse sweeps after the
/// coarse correction); apply is one forward sweep from zero on the host.
template <class T = real_t>
class gsGaussSeidelOp : public gsPreconditionerOp<T> {
public:
    explicit gsGaussSeidelOp(const gsSparseMatrix<T> &M) : m_M(&M) {}
    void apply(const gsMatrix<T> &input, gsMatrix<T> &x) const override
    {
        // column-major storage: solve (D + L) x = r column by column
        const index_t n = m_M->rows();
        x.resize(n, 1);
        for (index_t i = 0; i < n; ++i) x(i) = input(i);
        for (index_t j = 0; j < n; ++j) {
            x(j) /= m_M->coeff(j, j);
            for (index_t k = m_M->outerIndexPtr()[j]; k < m_M->outerIndexPtr()[j + 1]; ++k) {
                const index_t r = m_M->innerIndexPtr()[k];
                if (r > j) x(r) -= m_M->valuePtr()[k] * x(j);
            }
        }
    }
    index_t rows() const override { return m_M->rows(); }
    index_t cols() const override { return m_M->cols(); }
    int kind() const override { return GSX_PREC_NONE; }
    int smootherKind() const override { return GSX_SMOOTH_GAUSS_SEIDEL; }

private:
    const gsSparseMatrix<T> *m_M;
};

template <class T>
std::shared_ptr<gsGaussSeidelOp<T>> makeGaussSeidelOp(const gsSparseMatrix<T> &M)
{
    return std::make_shared<gsGaussSeidelOp<T>>(M);
}

/// Exact solvers of gsMatrixOp.h (makeSparseCholeskySolver /
/// makeSparseLUSolver, src/gsSolver/gsMatrixOp.h:306-370): apply solves with
/// a dense FP64 factorisation on the host.  As the coarse solver of a
/// gsMultiGridOp they select its exact coarsest-level solve, which runs on
/// the device (the dense FP64 inverse of the coarsest operator).
template <class T = real_t>
class gsExactSolverOp : public gsLinearOperator<T> {
public:
    gsExactSolverOp(const gsSparseMatrix<T> &M, bool cholesky) : m_n(M.rows()), m_chol(cholesky)
    {
        const index_t n = m_n;
        m_F.assign((size_t)n * n, 0.0);
        for (index_t j = 0; j < M.cols(); ++j)
            for (index_t k = M.outerIndexPtr()[j]; k < M.outerIndexPtr()[j + 1]; ++k)
                m_F[(size_t)M.innerIndexPtr()[k] * n + j] = M.valuePtr()[k];
        m_piv.resize((size_t)n);
        for (index_t c = 0; c < n; ++c) {
            index_t pr = c;
            if (!m_chol)
                for (index_t r = c + 1; r < n; ++r)
                    if (std::fabs(m_F[(size_t)r * n + c]) > std::fabs(m_F[(size_t)pr * n + c])) pr = r;
            if (m_F[(size_t)pr * n + c] == 0.0 || (m_chol && m_F[(size_t)c * n + c] <= 0.0))
                throw std::runtime_error(m_chol ? "makeSparseCholeskySolver: matrix not positive definite"
                                                : "makeSparseLUSolver: singular matrix");
            m_piv[(size_t)c] = pr;
            if (pr != c)
                for (index_t q = 0; q < n; ++q) std::swap(m_F[(size_t)c * n + q], m_F[(size_t)pr * n + q]);
            for (index_t r = c + 1; r < n; ++r) {
                const double l = m_F[(size_t)r * n + c] / m_F[(size_t)c * n + c];
                m_F[(size_t)r * n + c] = l;
                if (l != 0.0)
                    for (index_t q = c + 1; q < n; ++q) m_F[(size_t)r * n + q] -= l * m_F[(size_t)c * n + q];
            }
        }
    }
    void apply(const gsMatrix<T> &input, gsMatrix<T> &x) const override
    {
        const index_t n = m_n;
        x.resize(n, 1);
        for (index_t i = 0; i < n; ++i) x(i) = input(i);
        for (index_t c = 0; c < n; ++c)
            if (m_piv[(size_t)c] != c) std::swap(x(c), x(m_piv[(size_t)c]));
        for (index_t r = 0; r < n; ++r)
            for (index_t c = 0; c < r; ++c) x(r) -= m_F[(size_t)r * n + c] * x(c);
        for (index_t r = n - 1; r >= 0; --r) {
            for (index_t c = r + 1; c < n; ++c) x(r) -= m_F[(size_t)r * n + c] * x(c);
            x(r) /= m_F[(size_t)r * n + r];
        }
    }
    index_t rows() const override { return m_n; }
    index_t cols() const override { return m_n; }

private:
    index_t m_n;
    bool m_chol;
    std::vector<double> m_F;
    std::vector<index_t> m_piv;
};

template <class T>
typename gsLinearOperator<T>::Ptr makeSparseCholeskySolver(const gsSparseMatrix<T> &M)
{
    return std::make_shared<gsExactSolverOp<T>>(M, true);
}
template <class T>
typename gsLinearOperator<T>::Ptr makeSparseLUSolver(const gsSparseMatrix<T> &M)
{
    return std::make_shared<gsExactSolverOp<T>>(M, false);
}

/// gsMultiGridOp (src/gsMultiGrid/gsMultiGrid.h:57-395) on the device.
/// Two ways to build it:
///  - make(fineMatrix, transferMatrices[, coarseSolver]), the reference's
///    (gsMultiGrid.h:136-141): the caller's transfers (gsGridHierarchy) are
///    uploaded and the Galerkin products P^T A P formed on the device;
///  - make(fineMatrix, numLevels[, opt]): the hierarchy from coarsening the
///    knot vectors of the problem the matrix was assembled on (transfers
///    and Galerkin products built on the device, from the kept element
///    matrices when they fit).
/// Per level, setSmoother(i, makeJacobiOp / makeRichardsonOp /
/// makeGaussSeidelOp (mg->matrix(i))) selects that level's smoother;
/// without it the Smoother option applies to every level.  Options
/// (setOptions): NumPreSmooth, NumPostSmooth, CoarseGridCorrectionDamping,
/// Damping (default Jacobi/Richardson damping), Smoother ("Jacobi"/"j",
/// "GaussSeidel"/"gs", "Richardson"/"r", "Chebyshev"/"c",
/// "SubspaceCorrectedMassSmoother"/"scms": the example's components --
/// patch interiors and 3D faces with box smoothers damped by their power
/// iteration, sides, edges and corners solved exactly; several components
/// need "fp64"), ChebyshevRatio, Scaling (sigma
/// of the scms smoother, 0.12), Precision ("fp64" or "mixed": FP32 cycle).
/// The coarsest level is always solved exactly (dense FP64 inverse on the
/// device); setCoarseSolver accepts makeSparseCholeskySolver /
/// makeSparseLUSolver.
template <class T = real_t>
class gsMultiGridOp : public gsPreconditionerOp<T> {
public:
    typedef std::shared_ptr<gsMultiGridOp> Ptr;
    typedef std::shared_ptr<gsMultiGridOp> uPtr;
    typedef typename gsLinearOperator<T>::Ptr OpPtr;
    typedef typename gsPreconditionerOp<T>::Ptr PrecondPtr;

    static Ptr make(const gsSparseMatrix<T> &fine, index_t numLevels, const gsOptionList &opt = gsOptionList())
    {
        Ptr p(new gsMultiGridOp());
        p->m_n = fine.rows();
        p->m_A = fine.devicePtr();
        p->m_prob = fine.problemPtr();
        if (!p->m_A || !p->m_prob)
            throw std::runtime_error("gsMultiGridOp: the matrix must come from gsExprAssembler/gsPoissonAssembler");
        p->m_levels = numLevels;
        p->m_opt = opt;
        p->m_smoothers.assign((size_t)numLevels, nullptr);
        return p;
    }
    static Ptr make(const gsSparseMatrix<T> &fine, std::vector<gsSparseMatrix<T, RowMajor>> transferMatrices,
                    OpPtr coarseSolver = OpPtr())
    {
        Ptr p(new gsMultiGridOp());
        p->m_n = fine.rows();
        p->m_A = detail::deviceMatrix(fine, false);
        p->m_transfers = std::move(transferMatrices);
        p->m_byTransfers = true;
        p->m_levels = (index_t)p->m_transfers.size() + 1;
        for (size_t l = 0; l < p->m_transfers.size(); ++l) {
            const index_t nf = l + 1 < p->m_transfers.size() ? p->m_transfers[l + 1].cols() : p->m_n;
            if (p->m_transfers[l].rows() != nf)
                throw std::runtime_error("gsMultiGridOp: transfer matrix sizes do not fit the fine matrix");
        }
        p->m_smoothers.assign((size_t)p->m_levels, nullptr);
        if (coarseSolver) p->setCoarseSolver(coarseSolver);
        return p;
    }
    void setOptions(const gsOptionList &o) override
    {
        const bool precChanged = o.askString("Precision", "fp64") != m_opt.askString("Precision", "fp64");
        m_opt = o;
        if (!m_mg) return;
        if (precChanged || !m_byTransfers) {
            m_mg.reset();   // rebuilt with the new options on the next use
            return;
        }
        detail::check(gsx_mg_set_cycle(m_mg.get(), m_opt.askInt("NumPreSmooth", 1), m_opt.askInt("NumPostSmooth", 1),
                                       m_opt.askReal("CoarseGridCorrectionDamping", 1.0)),
                      "gsx_mg_set_cycle");
        applySmoothers(m_mg.get());
    }
    index_t numLevels() const { return m_levels; }
    /// the operator of level lvl (0 = coarsest), a host copy of the device's
    const gsSparseMatrix<T> &matrix(index_t lvl) const
    {
        if (lvl < 0 || lvl >= m_levels) throw std::runtime_error("gsMultiGridOp::matrix: level out of range");
        if (m_mats.empty()) m_mats.resize((size_t)m_levels);
        std::shared_ptr<gsSparseMatrix<T>> &M = m_mats[(size_t)lvl];
        if (!M) {
            gsx_mg *h = device();
            int64_t n = 0, nnz = 0, pn = 0;
            detail::check(gsx_mg_level_info(h, (int32_t)lvl, &n, &nnz, &pn), "gsx_mg_level_info");
            std::vector<int32_t> rp((size_t)n + 1), ci((size_t)nnz);
            std::vector<double> v((size_t)nnz);
            detail::check(gsx_mg_level_download(h, (int32_t)lvl, rp.data(), ci.data(), v.data(), nullptr, nullptr,
                                                nullptr),
                          "gsx_mg_level_download");
            gsSparseMatrix<T, RowMajor> R;
            R.setFromCompressed((index_t)n, (index_t)n, std::vector<index_t>(rp.begin(), rp.end()),
                                std::vector<index_t>(ci.begin(), ci.end()), std::move(v));
            // row-major -> column-major storage
            std::vector<typename gsSparseMatrix<T>::Triplet> t;
            t.reserve((size_t)nnz);
            for (index_t i = 0; i < (index_t)n; ++i)
                for (index_t k = R.outerIndexPtr()[i]; k < R.outerIndexPtr()[i + 1]; ++k)
                    t.push_back({i, R.innerIndexPtr()[k], R.valuePtr()[k]});
            M = std::make_shared<gsSparseMatrix<T>>();
            M->setFromTriplets((index_t)n, (index_t)n, t, false);
        }
        return *M;
    }
    OpPtr underlyingOp(index_t lvl) const { return makeMatrixOp(matrix(lvl)); }
    void setSmoother(index_t lvl, const PrecondPtr &sm)
    {
        if (lvl < 1 || lvl >= m_levels) throw std::runtime_error("gsMultiGridOp::setSmoother: level out of range");
        if (!sm || sm->smootherKind() < 0)
            throw std::runtime_error("gsMultiGridOp::setSmoother: Jacobi, Richardson or Gauss-Seidel smoothers only");
        m_smoothers[(size_t)lvl] = sm;
        if (m_mg) applySmoothers(m_mg.get());
    }
    PrecondPtr smoother(index_t lvl) const { return m_smoothers.at((size_t)lvl); }
    void setCoarseSolver(const OpPtr &op)
    {
        if (!std::dynamic_pointer_cast<gsExactSolverOp<T>>(op))
            throw std::runtime_error("gsMultiGridOp::setCoarseSolver: exact solvers (makeSparseCholeskySolver, "
                                     "makeSparseLUSolver) only");
        m_coarse = op;
    }
    void apply(const gsMatrix<T> &input, gsMatrix<T> &x) const override
    {
        detail::DeviceVec a(sizeof(double) * m_n), b(sizeof(double) * m_n);
        detail::check(gsx_copy_to_device(detail::ctx(), a.p, input.data(), sizeof(double) * m_n), "upload");
        detail::check(gsx_mg_apply(device(), a.d(), b.d()), "gsx_mg_apply");
        x.resize(m_n, 1);
        detail::check(gsx_copy_to_host(detail::ctx(), x.data(), b.p, sizeof(double) * m_n), "download");
    }
    index_t rows() const override { return m_n; }
    index_t cols() const override { return m_n; }
    int kind() const override { return GSX_PREC_MG; }
    gsx_mg *device() const
    {
        if (!m_mg) {
            gsx_mg *h = nullptr;
            const int prec = m_opt.askString("Precision", "fp64") == "mixed" ? GSX_MIXED : GSX_FP64;
            const int npre = m_opt.askInt("NumPreSmooth", 1), npost = m_opt.askInt("NumPostSmooth", 1);
            const double tau = m_opt.askReal("Damping", 1.0), cgc = m_opt.askReal("CoarseGridCorrectionDamping", 1.0);
            if (m_byTransfers) {
                const int nt = (int)m_transfers.size();
                std::vector<std::vector<int32_t>> rp(nt), ci(nt);
                std::vector<std::vector<double>> v(nt);
                std::vector<const int32_t *> prp(nt), pci(nt);
                std::vector<const double *> pv(nt);
                std::vector<int32_t> nc(nt);
                for (int l = 0; l < nt; ++l) {
                    m_transfers[(size_t)l].toCsr(rp[l], ci[l], v[l]);
                    prp[l] = rp[l].data();
                    pci[l] = ci[l].data();
                    pv[l] = v[l].data();
                    nc[l] = m_transfers[(size_t)l].cols();
                }
                std::shared_ptr<gsx_matrix> A = m_A;
                detail::check(gsx_mg_create_transfers(A.get(), nt, nc.data(), prp.data(), pci.data(), pv.data(), npre,
                                                      npost, 1, tau, cgc, prec, &h),
                              "gsx_mg_create_transfers");
            } else {
                detail::check(gsx_mg_create(m_prob.get(), m_A.get(), m_levels, npre, npost, 1, tau, cgc, prec, &h),
                              "gsx_mg_create");
            }
            m_mg.reset(h, [](gsx_mg *q) { gsx_mg_destroy(q); });
            // multiGrid_example's MG.Smoother names (r, j, gs, scms) plus Chebyshev (c)
            const std::string sm = m_opt.askString("Smoother", "Jacobi");
            int kind = -1;
            if (sm == "Jacobi" || sm == "j") kind = GSX_SMOOTH_JACOBI;
            else if (sm == "GaussSeidel" || sm == "gs") kind = GSX_SMOOTH_GAUSS_SEIDEL;
            else if (sm == "Richardson" || sm == "r") kind = GSX_SMOOTH_RICHARDSON;
            else if (sm == "Chebyshev" || sm == "c") kind = GSX_SMOOTH_CHEBYSHEV;
            else if (sm == "SubspaceCorrectedMassSmoother" || sm == "scms") kind = GSX_SMOOTH_SCMS;
            if (kind < 0) throw std::runtime_error("gsMultiGridOp: smoother '" + sm + "' is not supported");
            bool perLevel = false;
            for (auto &q : m_smoothers) perLevel = perLevel || q;
            if (!perLevel) {
                const double param = kind == GSX_SMOOTH_SCMS ? m_opt.askReal("Scaling", 0.12)
                                                             : m_opt.askReal("ChebyshevRatio", 30.0);
                detail::check(gsx_mg_set_smoother(h, kind, param), "gsx_mg_set_smoother");
            }
            applySmoothers(h);
        }
        return m_mg.get();
    }

private:
    void applySmoothers(gsx_mg *h) const
    {
        for (index_t l = 1; l < m_levels; ++l) {
            const PrecondPtr &q = m_smoothers[(size_t)l];
            if (!q) continue;
            detail::check(gsx_mg_set_level_smoother(h, (int32_t)l, q->smootherKind(), (double)q->smootherDamping(), 0.0),
                          "gsx_mg_set_level_smoother");
        }
    }
    index_t m_n = 0;
    index_t m_levels = 1;
    bool m_byTransfers = false;
    gsOptionList m_opt;
    std::shared_ptr<gsx_matrix> m_A;
    std::shared_ptr<gsx_problem> m_prob;
    std::vector<gsSparseMatrix<T, RowMajor>> m_transfers;
    std::vector<PrecondPtr> m_smoothers;
    OpPtr m_coarse;
    mutable std::vector<std::shared_ptr<gsSparseMatrix<T>>> m_mats;
    mutable std::shared_ptr<gsx_mg> m_mg;
};

/// The operator of gsPatchPreconditionersCreator::fastDiagonalizationOp
/// (src/gsSolver/gsPatchPreconditionersCreator.hpp:258-348): the exact
/// inverse of alpha M + beta sum_k M x .. x K_k x .. x M on the parameter
/// domain of one patch, built on the device (gsx_fd_create).
template <class T = real_t>
class gsFastDiagonalizationOp : public gsPreconditionerOp<T> {
public:
    gsFastDiagonalizationOp(const gsTensorBSplineBasis<T> &basis, const gsBoundaryConditions<T> &bc, T alpha, T beta)
        : gsFastDiagonalizationOp(basis, bc, alpha, beta, T(0))
    {
    }

protected:
    // sigma > 0: the subspace corrected mass smoother of the same basis
    gsFastDiagonalizationOp(const gsTensorBSplineBasis<T> &basis, const gsBoundaryConditions<T> &bc, T alpha, T beta,
                            T sigma)
    {
        const int d = basis.dim();
        if (d < 2 || d > 3) throw std::runtime_error("fastDiagonalizationOp: 2D and 3D bases only");
        gsx_patch_desc D;
        std::memset(&D, 0, sizeof(D));
        D.d = d;
        int nb[3] = {1, 1, 1};
        std::vector<std::vector<double>> gr(d);
        for (int j = 0; j < d; ++j) {
            D.p[j] = basis.degree(j);
            D.nknots[j] = (int32_t)basis.knots(j).size();
            D.knots_host[j] = basis.knots(j).knots().data();
            nb[j] = (int)basis.knots(j).numBasis();
            // Greville abscissae: the identity map (the operator does not
            // depend on the geometry)
            const auto &U = basis.knots(j).knots();
            const int p = D.p[j];
            for (int i = 0; i < nb[j]; ++i) {
                double g = 0.0;
                for (int k = 1; k <= p; ++k) g += U[i + k];
                gr[j].push_back(p > 0 ? g / p : U[i]);
            }
        }
        const index_t n = (index_t)nb[0] * nb[1] * nb[2];
        std::vector<double> ctrl((size_t)n * d);
        for (index_t i = 0; i < n; ++i) {
            const int c[3] = {(int)(i % nb[0]), (int)((i / nb[0]) % nb[1]), (int)(i / ((index_t)nb[0] * nb[1]))};
            for (int j = 0; j < d; ++j) ctrl[(size_t)i * d + j] = gr[j][c[j]];
        }
        D.ctrl_host = ctrl.data();
        D.fkind = 1;
        std::vector<int32_t> ds;
        for (auto &s : bc.dirichletSides())
            if (s.first == 0) { ds.push_back(0); ds.push_back(s.second); }
        gsx_problem *p = nullptr;
        detail::check(gsx_problem_create(detail::ctx(), 1, &D, (int32_t)ds.size() / 2, ds.empty() ? nullptr : ds.data(),
                                         0, nullptr, &p),
                      "gsx_problem_create");
        m_prob.reset(p, detail::ProblemDeleter());
        int64_t nbasis = 0, nfree = 0, nelem = 0, nglued = 0;
        detail::check(gsx_problem_info(p, &nbasis, &nfree, &nelem, &nglued), "gsx_problem_info");
        m_n = (index_t)nfree;
        gsx_fd *f = nullptr;
        if (sigma > 0)
            detail::check(gsx_scms_create(p, sigma, alpha, beta, &f), "gsx_scms_create");
        else
            detail::check(gsx_fd_create(p, alpha, beta, &f), "gsx_fd_create");
        m_fd.reset(f, [](gsx_fd *q) { gsx_fd_destroy(q); });
    }

public:
    void apply(const gsMatrix<T> &input, gsMatrix<T> &x) const override
    {
        detail::DeviceVec a(sizeof(double) * std::max<index_t>(1, m_n)), b(sizeof(double) * std::max<index_t>(1, m_n));
        detail::check(gsx_copy_to_device(detail::ctx(), a.p, input.data(), sizeof(double) * m_n), "upload");
        detail::check(gsx_fd_apply(m_fd.get(), a.d(), b.d()), "gsx_fd_apply");
        x.resize(m_n, 1);
        detail::check(gsx_copy_to_host(detail::ctx(), x.data(), b.p, sizeof(double) * m_n), "download");
    }
    index_t rows() const override { return m_n; }
    index_t cols() const override { return m_n; }
    int kind() const override { return GSX_PREC_FD; }
    gsx_fd *device() const { return m_fd.get(); }

private:
    index_t m_n = 0;
    std::shared_ptr<gsx_problem> m_prob;
    std::shared_ptr<gsx_fd> m_fd;
};

/// The operator of gsPatchPreconditionersCreator::subspaceCorrectedMassSmootherOp
/// (src/gsSolver/gsPatchPreconditionersCreator.hpp:574-758): the sum over the
/// 2^d combinations of each direction's tilde space (vanishing odd
/// derivatives at the ends) and its complement, built on the device
/// (gsx_scms_create); CG uses it like the fast diagonalization.
template <class T = real_t>
class gsSubspaceCorrectedMassSmootherOp : public gsFastDiagonalizationOp<T> {
public:
    gsSubspaceCorrectedMassSmootherOp(const gsTensorBSplineBasis<T> &basis, const gsBoundaryConditions<T> &bc,
                                      T sigma, T alpha, T beta)
        : gsFastDiagonalizationOp<T>(basis, bc, alpha, beta, check_sigma(sigma))
    {
    }

private:
    static T check_sigma(T sigma)
    {
        if (!(sigma > 0)) throw std::runtime_error("subspaceCorrectedMassSmootherOp: sigma must be positive");
        return sigma;
    }
};

/// gsPatchPreconditionersCreator (src/gsSolver/gsPatchPreconditionersCreator.h)
/// subset: fastDiagonalizationOp and subspaceCorrectedMassSmootherOp with
/// DirichletStrategy elimination (gamma = 0).
template <class T = real_t>
struct gsPatchPreconditionersCreator {
    typedef std::shared_ptr<gsPreconditionerOp<T>> OpUPtr;
    static gsOptionList defaultOptions()
    {
        gsOptionList o;
        o.setInt("DirichletStrategy", dirichlet::elimination);
        return o;
    }
    static OpUPtr fastDiagonalizationOp(const gsTensorBSplineBasis<T> &basis, const gsBoundaryConditions<T> &bc,
                                        const gsOptionList &opt = defaultOptions(), T alpha = 0, T beta = 1,
                                        T gamma = 0)
    {
        if (opt.askInt("DirichletStrategy", dirichlet::elimination) != dirichlet::elimination)
            throw std::runtime_error("fastDiagonalizationOp: Dirichlet elimination only");
        if (gamma != 0) throw std::runtime_error("fastDiagonalizationOp: gamma != 0 (pure Neumann averaging) is not provided");
        return std::make_shared<gsFastDiagonalizationOp<T>>(basis, bc, alpha, beta);
    }
    /// Defaults sigma = 12/100, alpha = 0, beta = 1 (gsPatchPreconditionersCreator.h:125-150).
    static OpUPtr subspaceCorrectedMassSmootherOp(const gsTensorBSplineBasis<T> &basis,
                                                  const gsBoundaryConditions<T> &bc,
                                                  const gsOptionList &opt = defaultOptions(),
                                                  T sigma = (T)(12) / (T)(100), T alpha = 0, T beta = 1)
    {
        if (opt.askInt("DirichletStrategy", dirichlet::elimination) != dirichlet::elimination)
            throw std::runtime_error("subspaceCorrectedMassSmootherOp: Dirichlet elimination only");
        if (beta == 0) throw std::runtime_error("subspaceCorrectedMassSmootherOp: beta == 0");
        return std::make_shared<gsSubspaceCorrectedMassSmootherOp<T>>(basis, bc, sigma, alpha, beta);
    }
};

/// gsIterativeSolver / gsConjugateGradient (src/gsSolver/gsIterativeSolver.h:
/// 20-230, src/gsSolver/gsConjugateGradient.hpp:20-93).  Options:
/// MaxIterations (1000), Tolerance (1e-10), Precision ("fp64" or "mixed").
template <class T = real_t>
class gsIterativeSolver {
public:
    gsIterativeSolver(const gsSparseMatrix<T> &A, typename gsPreconditionerOp<T>::Ptr precond = nullptr)
        : m_A(&A), m_prec(std::move(precond)) {}
    virtual ~gsIterativeSolver() = default;
    gsIterativeSolver &setOptions(const gsOptionList &o)
    {
        m_max = o.askInt("MaxIterations", m_max);
        m_tol = o.askReal("Tolerance", m_tol);
        m_precision = o.askString("Precision", m_precision);
        return *this;
    }
    gsIterativeSolver &setTolerance(T t) { m_tol = t; return *this; }
    gsIterativeSolver &setMaxIterations(index_t m) { m_max = m; return *this; }
    T tolerance() const { return m_tol; }
    index_t maxIterations() const { return m_max; }
    index_t iterations() const { return m_it; }
    T error() const { return m_err; }
    void solve(const gsMatrix<T> &rhs, gsMatrix<T> &x) { run(rhs, x, nullptr); }
    void solveDetailed(const gsMatrix<T> &rhs, gsMatrix<T> &x, gsMatrix<T> &hist) { run(rhs, x, &hist); }

protected:
    void run(const gsMatrix<T> &rhs, gsMatrix<T> &x, gsMatrix<T> *hist)
    {
        const index_t n = m_A->rows();
        const bool mixed = m_precision == "mixed";
        std::shared_ptr<gsx_matrix> dev = detail::deviceMatrix(*m_A, mixed);
        if (x.size() == 0) x.setZero(n, 1);
        detail::DeviceVec b(sizeof(double) * n), xd(sizeof(double) * n);
        detail::check(gsx_copy_to_device(detail::ctx(), b.p, rhs.data(), sizeof(double) * n), "upload rhs");
        detail::check(gsx_copy_to_device(detail::ctx(), xd.p, x.data(), sizeof(double) * n), "upload x");
        int32_t it = 0;
        double rel = 0;
        int kind = m_prec ? m_prec->kind() : GSX_PREC_NONE;
        double tau = 1.0;
        gsx_mg *mg = nullptr;
        if (kind == GSX_PREC_JACOBI) tau = static_cast<gsJacobiOp<T> *>(m_prec.get())->damping();
        if (kind == GSX_PREC_MG) mg = static_cast<gsMultiGridOp<T> *>(m_prec.get())->device();
        const gsx_status st =
            kind == GSX_PREC_FD
                ? gsx_cg_solve_fd(dev.get(), static_cast<gsFastDiagonalizationOp<T> *>(m_prec.get())->device(),
                                  mixed ? GSX_MIXED : GSX_FP64, b.d(), xd.d(), m_tol, m_max, &it, &rel)
                : gsx_cg_solve(dev.get(), kind, tau, mg, mixed ? GSX_MIXED : GSX_FP64, b.d(), xd.d(), m_tol, m_max,
                               &it, &rel);
        if (st != GSX_OK && st != GSX_ERR_NOT_CONVERGED) detail::check(st, "gsx_cg_solve");
        m_it = it;
        m_err = rel;
        detail::check(gsx_copy_to_host(detail::ctx(), x.data(), xd.p, sizeof(double) * n), "download x");
        if (hist) {
            int32_t cnt = 0;
            gsx_cg_history(dev.get(), nullptr, 0, &cnt);
            hist->resize(cnt, 1);
            gsx_cg_history(dev.get(), hist->data(), cnt, &cnt);
        }
    }
    const gsSparseMatrix<T> *m_A;
    typename gsPreconditionerOp<T>::Ptr m_prec;
    index_t m_max = 1000;
    T m_tol = 1e-10;
    std::string m_precision = "fp64";
    index_t m_it = 0;
    T m_err = 0;
};

template <class T = real_t>
class gsConjugateGradient : public gsIterativeSolver<T> {
public:
    gsConjugateGradient(const gsSparseMatrix<T> &A, typename gsPreconditionerOp<T>::Ptr precond = nullptr)
        : gsIterativeSolver<T>(A, std::move(precond)) {}
    static gsOptionList defaultOptions()
    {
        gsOptionList o;
        o.addInt("MaxIterations", "", 1000);
        o.addReal("Tolerance", "", 1e-10);
        o.addString("Precision", "", "fp64");
        return o;
    }
};

// ---------------------------------------------------------------- evaluation
/// gsExprEvaluator (src/gsAssembler/gsExprEvaluator.h:38,382-450) for the
/// error of a discrete solution against u = prod sin(pi x_i), the two
/// integrals of examples/poisson2_example.cpp:173-176:
///   math::sqrt(ev.integral((u_ex - u_sol).sqNorm() * meas(G)))
///   math::sqrt(ev.integral((igrad(u_ex) - igrad(u_sol, G)).sqNorm() * meas(G)))
/// (fixed Dirichlet values included once boundary data is set)
template <class T = real_t>
class gsExprEvaluator {
public:
    explicit gsExprEvaluator(const gsExprAssembler<T> &A) : m_ea(&A) {}
    explicit gsExprEvaluator(const gsPoissonAssembler<T> &A) : m_pa(&A) {}
    expr::exact_var getVariable(const gsFunctionExpr<T> &f, const geometryMap & = geometryMap()) const
    {
        if (f.kind() != 3)
            throw std::runtime_error("gsExprEvaluator: only the manufactured solution prod sin(pi x_i) is evaluated");
        return expr::exact_var();
    }
    T integral(const expr::l2_integrand &e) const
    {
        const T err = l2ErrorSinProduct(*e.d.coefs);
        return err * err;
    }
    T integral(const expr::h1_integrand &e) const
    {
        T l2 = 0, semi = 0;
        errorNorms(*e.d.coefs, l2, semi);
        return semi * semi;
    }
    T l2ErrorSinProduct(const gsMatrix<T> &sol) const
    {
        T l2 = 0, semi = 0;
        errorNorms(sol, l2, semi);
        return l2;
    }

private:
    void errorNorms(const gsMatrix<T> &sol, T &l2, T &semi) const
    {
        std::shared_ptr<gsx_problem> prob = m_ea ? m_ea->problem() : m_pa->problem();
        if (!prob) throw std::runtime_error("gsExprEvaluator: nothing assembled yet");
        int64_t nbasis = 0, nfree = 0, nelem = 0, nglued = 0;
        detail::check(gsx_problem_info(prob.get(), &nbasis, &nfree, &nelem, &nglued), "gsx_problem_info");
        if (sol.rows() != nfree) throw std::runtime_error("gsExprEvaluator: solution size does not match the space");
        detail::DeviceVec s(sizeof(double) * std::max<int64_t>(1, nfree));
        detail::check(gsx_copy_to_device(detail::ctx(), s.p, sol.data(), sizeof(double) * nfree), "upload");
        double e = 0, h = 0;
        detail::check(gsx_error_norms(prob.get(), s.d(), &e, &h), "gsx_error_norms");
        l2 = e;
        semi = h;
    }

    const gsExprAssembler<T> *m_ea = nullptr;
    const gsPoissonAssembler<T> *m_pa = nullptr;
};

namespace math {
using std::sqrt;
using std::log;
using std::abs;
}

/// gsSparseSolver<T>::CGDiagonal: Eigen's diagonally preconditioned CG
/// (tolerance machine epsilon, at most 2n iterations) on the device.
template <class T = real_t>
struct gsSparseSolver {
    class CGDiagonal {
    public:
        CGDiagonal &compute(const gsSparseMatrix<T> &A)
        {
            m_A = A;
            m_prec = makeJacobiOp(m_A, T(1));
            return *this;
        }
        gsMatrix<T> solve(const gsMatrix<T> &b)
        {
            gsConjugateGradient<T> cg(m_A, m_prec);
            cg.setTolerance(m_tol).setMaxIterations(m_max > 0 ? m_max : 2 * m_A.rows());
            gsMatrix<T> x;
            x.setZero(b.rows(), 1);
            cg.solve(b, x);
            m_it = cg.iterations();
            m_err = cg.error();
            return x;
        }
        void setTolerance(T t) { m_tol = t; }
        void setMaxIterations(index_t m) { m_max = m; }
        index_t iterations() const { return m_it; }
        T error() const { return m_err; }

    private:
        gsSparseMatrix<T> m_A;
        std::shared_ptr<gsJacobiOp<T>> m_prec;
        T m_tol = std::numeric_limits<T>::epsilon();
        index_t m_max = 0, m_it = 0;
        T m_err = 0;
    };
};

}  // namespace gismo
