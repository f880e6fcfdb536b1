// gsXml.h -- G+Smo XML files for the drop-in classes of gismo.h.
//
// Reads and writes the reference's XML format (src/gsIO/gsFileData.h:33-190,
// src/gsIO/gsXmlUtils.hpp, src/gsIO/gsXml.hpp, src/gsIO/gsXml.cpp) for the
// objects of the Poisson path:
//   <Geometry type="TensorBSpline2|TensorBSpline3" id=..> with a
//     <Basis type="TensorBSplineBasisD"> of <Basis type="BSplineBasis"
//     index=k><KnotVector degree=p>..</KnotVector></Basis> and
//     <coefs geoDim=..> (row-major control points)
//   <MultiPatch parDim=.. id=..> with <patches type="id_range|id_index">,
//     <interfaces> (p1 s1 p2 s2, direction map [d], orientation [d]) and
//     <boundary name=..> (patch side) lists
//   <Function type="FunctionExpr" dim=.. id=..>expression</Function>, or
//     vector-valued with <c index=..> components
//   <OptionList id=..> with <int|real|bool|string label=.. value=..>
//   <boundaryConditions id=.. multipatch=..> with <bc type="Dirichlet|
//     Neumann" name=.. function=..> referring to the MultiPatch's named
//     boundaries
//   <Matrix rows cols> (row-major values), <SparseMatrix rows cols>
//     (row col value per non-zero, column by column)
// Rational (NURBS) geometries, THB bases and base64 payloads are refused
// with std::runtime_error.
#pragma once

#include "gismo.h"

#include <cctype>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace gismo {

namespace xml {

/// One element of an XML document.
struct Node {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string text;              // concatenated character data
    std::vector<Node> children;

    const std::string *attr(const std::string &k) const
    {
        for (auto &a : attrs)
            if (a.first == k) return &a.second;
        return nullptr;
    }
    std::string attr_or(const std::string &k, const std::string &def = "") const
    {
        const std::string *v = attr(k);
        return v ? *v : def;
    }
    const Node *child(const std::string &n) const
    {
        for (auto &c : children)
            if (c.name == n) return &c;
        return nullptr;
    }
};

class Parser {
public:
    explicit Parser(const std::string &src) : s(src) {}

    Node document()
    {
        Node root;
        root.name = "#document";
        parse_content(root, true);
        return root;
    }

private:
    const std::string &s;
    size_t i = 0;

    [[noreturn]] void error(const std::string &what) const
    {
        std::ostringstream os;
        os << "XML parse error at offset " << i << ": " << what;
        throw std::runtime_error(os.str());
    }
    bool starts(const char *t) const { return s.compare(i, std::strlen(t), t) == 0; }
    void skip_until(const char *t)
    {
        const size_t j = s.find(t, i);
        if (j == std::string::npos) error(std::string("missing ") + t);
        i = j + std::strlen(t);
    }
    static std::string decode(const std::string &t)
    {
        std::string out;
        for (size_t k = 0; k < t.size(); ++k) {
            if (t[k] != '&') {
                out += t[k];
                continue;
            }
            const size_t e = t.find(';', k);
            const std::string ent = t.substr(k + 1, e - k - 1);
            if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "amp") out += '&';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else out += "&" + ent + ";";
            k = e;
        }
        return out;
    }
    std::string name()
    {
        const size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' ||
                                s[i] == '.'))
            ++i;
        if (i == b) error("name expected");
        return s.substr(b, i - b);
    }
    void ws()
    {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    void parse_content(Node &parent, bool top)
    {
        while (i < s.size()) {
            if (starts("<?")) {
                skip_until("?>");
            } else if (starts("<!--")) {
                skip_until("-->");
            } else if (starts("<![CDATA[")) {
                i += 9;
                const size_t e = s.find("]]>", i);
                if (e == std::string::npos) error("unterminated CDATA");
                parent.text += s.substr(i, e - i);
                i = e + 3;
            } else if (starts("<!")) {
                skip_until(">");
            } else if (starts("</")) {
                if (top) error("unexpected closing tag");
                i += 2;
                const std::string n = name();
                if (n != parent.name) error("mismatched </" + n + "> for <" + parent.name + ">");
                ws();
                if (i >= s.size() || s[i] != '>') error("'>' expected");
                ++i;
                return;
            } else if (s[i] == '<') {
                ++i;
                Node c;
                c.name = name();
                bool empty = false;
                while (true) {
                    ws();
                    if (starts("/>")) {
                        i += 2;
                        empty = true;
                        break;
                    }
                    if (i < s.size() && s[i] == '>') {
                        ++i;
                        break;
                    }
                    const std::string k = name();
                    ws();
                    if (i >= s.size() || s[i] != '=') error("'=' expected");
                    ++i;
                    ws();
                    const char q = s[i];
                    if (q != '"' && q != '\'') error("quote expected");
                    const size_t e = s.find(q, i + 1);
                    if (e == std::string::npos) error("unterminated attribute");
                    c.attrs.push_back({k, decode(s.substr(i + 1, e - i - 1))});
                    i = e + 1;
                }
                if (!empty) parse_content(c, false);
                parent.children.push_back(std::move(c));
            } else {
                const size_t e = s.find('<', i);
                const size_t stop = e == std::string::npos ? s.size() : e;
                parent.text += decode(s.substr(i, stop - i));
                i = stop;
            }
        }
        if (!top) error("unterminated <" + parent.name + ">");
    }
};

inline std::vector<double> numbers(const std::string &t)
{
    std::vector<double> out;
    std::istringstream is(t);
    std::string tok;
    while (is >> tok) {
        char *end = nullptr;
        const double v = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str() || *end) throw std::runtime_error("XML: not a number: '" + tok + "'");
        out.push_back(v);
    }
    return out;
}

inline std::vector<long> integers(const std::string &t)
{
    std::vector<long> out;
    for (double v : numbers(t)) {
        if (v != std::floor(v)) throw std::runtime_error("XML: integer expected");
        out.push_back((long)v);
    }
    return out;
}

inline std::string escape(const std::string &t)
{
    std::string o;
    for (char c : t) {
        if (c == '<') o += "&lt;";
        else if (c == '>') o += "&gt;";
        else if (c == '&') o += "&amp;";
        else if (c == '"') o += "&quot;";
        else o += c;
    }
    return o;
}

}  // namespace xml

/// gsFileData (src/gsIO/gsFileData.h): an XML document of G+Smo objects.
template <class T = real_t>
class gsFileData {
public:
    gsFileData() = default;
    explicit gsFileData(const std::string &fn) { read(fn); }

    void read(const std::string &fn)
    {
        std::ifstream f(fn, std::ios::binary);
        if (!f) throw std::runtime_error("gsFileData: cannot open " + fn);
        std::stringstream ss;
        ss << f.rdbuf();
        readString(ss.str());
        m_path = fn;
    }
    void readString(const std::string &src)
    {
        xml::Parser p(src);
        xml::Node doc = p.document();
        const xml::Node *root = doc.child("xml");
        if (!root) throw std::runtime_error("gsFileData: no <xml> root element");
        m_root = *root;
    }
    const std::string &lastPath() const { return m_path; }
    const xml::Node &root() const { return m_root; }

    /// element with attribute id == id (any tag)
    const xml::Node &byId(int id) const
    {
        for (auto &c : m_root.children) {
            const std::string *v = c.attr("id");
            if (v && std::atoi(v->c_str()) == id) return c;
        }
        throw std::runtime_error("gsFileData: no object with id " + std::to_string(id));
    }
    const xml::Node &first(const std::string &tag) const
    {
        for (auto &c : m_root.children)
            if (c.name == tag) return c;
        throw std::runtime_error("gsFileData: no <" + tag + "> object");
    }
    bool has(const std::string &tag) const
    {
        for (auto &c : m_root.children)
            if (c.name == tag) return true;
        return false;
    }

    // ---- readers (gsFileData::getId / getFirst) ----------------------------
    void getId(int id, gsTensorBSplinePatch &g) const { g = geometry(byId(id)); }
    void getId(int id, gsMultiPatch<T> &mp) const { mp = multiPatch(byId(id)); }
    void getId(int id, gsFunctionExpr<T> &f) const { f = function(byId(id)); }
    void getId(int id, gsOptionList &o) const { o = options(byId(id)); }
    void getId(int id, gsMatrix<T> &m) const { m = matrix(byId(id)); }
    void getId(int id, gsSparseMatrix<T> &m) const { m = sparseMatrix(byId(id)); }
    void getId(int id, gsBoundaryConditions<T> &bc) const { bc = conditions(byId(id)); }

    void getFirst(gsMultiPatch<T> &mp) const
    {
        if (has("MultiPatch")) mp = multiPatch(first("MultiPatch"));
        else {   // a lone geometry becomes a one-patch multipatch
            mp = gsMultiPatch<T>();
            mp.addPatch(geometry(first("Geometry")));
            mp.computeTopology();
        }
    }
    void getFirst(gsTensorBSplinePatch &g) const { g = geometry(first("Geometry")); }
    void getFirst(gsMatrix<T> &m) const { m = matrix(first("Matrix")); }
    void getFirst(gsSparseMatrix<T> &m) const { m = sparseMatrix(first("SparseMatrix")); }

    // ---- writers (gsFileData::add / save) --------------------------------
    void add(const gsMatrix<T> &m, int id = -1)
    {
        std::ostringstream os;
        os << std::setprecision(std::numeric_limits<T>::max_digits10);
        for (index_t i = 0; i < m.rows(); ++i) {
            for (index_t j = 0; j < m.cols(); ++j) os << m(i, j) << " ";
            os << "\n";
        }
        push("Matrix", {{"rows", std::to_string(m.rows())}, {"cols", std::to_string(m.cols())}}, os.str(), id);
    }
    void add(const gsSparseMatrix<T> &m, int id = -1)
    {
        std::ostringstream os;
        os << std::setprecision(std::numeric_limits<T>::max_digits10);
        for (index_t j = 0; j < m.cols(); ++j)
            for (index_t k = m.outerIndexPtr()[j]; k < m.outerIndexPtr()[j + 1]; ++k)
                os << m.innerIndexPtr()[k] << " " << j << " " << m.valuePtr()[k] << "\n";
        push("SparseMatrix", {{"rows", std::to_string(m.rows())}, {"cols", std::to_string(m.cols())}}, os.str(), id);
    }
    void add(const gsTensorBSplinePatch &g, int id = -1)
    {
        xml::Node n;
        n.name = "Geometry";
        const int d = g.domainDim();
        n.attrs = {{"type", "TensorBSpline" + std::to_string(d)}};
        xml::Node basis;
        basis.name = "Basis";
        basis.attrs = {{"type", "TensorBSplineBasis" + std::to_string(d)}};
        for (int k = 0; k < d; ++k) {
            xml::Node b, kv;
            b.name = "Basis";
            b.attrs = {{"type", "BSplineBasis"}, {"index", std::to_string(k)}};
            kv.name = "KnotVector";
            kv.attrs = {{"degree", std::to_string(g.knots(k).degree())}};
            std::ostringstream os;
            os << std::setprecision(std::numeric_limits<T>::max_digits10);
            for (size_t i = 0; i < g.knots(k).size(); ++i) os << g.knots(k)[i] << " ";
            kv.text = os.str();
            b.children.push_back(kv);
            basis.children.push_back(b);
        }
        xml::Node coefs;
        coefs.name = "coefs";
        coefs.attrs = {{"geoDim", std::to_string(g.targetDim())}};
        std::ostringstream os;
        os << std::setprecision(std::numeric_limits<T>::max_digits10);
        for (index_t i = 0; i < g.coefs().rows(); ++i) {
            for (index_t r = 0; r < g.coefs().cols(); ++r) os << g.coefs()(i, r) << " ";
            os << "\n";
        }
        coefs.text = os.str();
        n.children = {basis, coefs};
        pushNode(n, id);
    }
    void add(const gsMultiPatch<T> &mp, int id = -1)
    {
        const int first_id = maxId() + 1;
        for (index_t k = 0; k < mp.nPatches(); ++k) add(mp.patch(k), first_id + (int)k);
        xml::Node n;
        n.name = "MultiPatch";
        n.attrs = {{"parDim", std::to_string(mp.dim())}};
        xml::Node patches;
        patches.name = "patches";
        patches.attrs = {{"type", "id_range"}};
        patches.text = std::to_string(first_id) + " " + std::to_string(first_id + mp.nPatches() - 1);
        n.children.push_back(patches);
        const int d = mp.dim();
        if (!mp.interfaces().empty()) {
            xml::Node ifs;
            ifs.name = "interfaces";
            std::ostringstream os;
            for (auto &f : mp.interfaces()) {
                // direction map and orientation over all d directions
                std::vector<int> map(d), orient(d, 1);
                const int dir1 = (f.s1 - 1) / 2, dir2 = (f.s2 - 1) / 2;
                map[dir1] = dir2;
                int t = 0;
                const int m[2] = {f.m0, f.m1}, fl[2] = {f.f0, f.f1};
                for (int j = 0; j < d; ++j) {
                    if (j == dir1) continue;
                    map[j] = m[t];
                    orient[j] = fl[t] ? 0 : 1;
                    ++t;
                }
                // the normal directions point against each other when the
                // sides have the same parity (west meets west)
                orient[dir1] = ((f.s1 - 1) % 2) != ((f.s2 - 1) % 2) ? 1 : 0;
                os << first_id + f.p1 << " " << f.s1 << " " << first_id + f.p2 << " " << f.s2;
                for (int j = 0; j < d; ++j) os << " " << map[j];
                for (int j = 0; j < d; ++j) os << " " << orient[j];
                os << "\n";
            }
            ifs.text = os.str();
            n.children.push_back(ifs);
        }
        if (!mp.boundaries().empty()) {
            xml::Node b;
            b.name = "boundary";
            std::ostringstream os;
            for (auto &s : mp.boundaries()) os << first_id + s.first << " " << s.second << "\n";
            b.text = os.str();
            n.children.push_back(b);
        }
        pushNode(n, id);
    }
    void save(const std::string &fn) const
    {
        std::ofstream f(fn);
        if (!f) throw std::runtime_error("gsFileData: cannot write " + fn);
        f << "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<xml>\n";
        for (auto &c : m_root.children) write(f, c, 1);
        f << "</xml>\n";
    }
    int maxId() const
    {
        int m = -1;
        for (auto &c : m_root.children) {
            const std::string *v = c.attr("id");
            if (v) m = std::max(m, std::atoi(v->c_str()));
        }
        return m;
    }

private:
    xml::Node m_root = [] {
        xml::Node n;
        n.name = "xml";
        return n;
    }();
    std::string m_path;

    void push(const std::string &tag, std::vector<std::pair<std::string, std::string>> attrs, const std::string &text,
              int id)
    {
        xml::Node n;
        n.name = tag;
        n.attrs = std::move(attrs);
        n.text = text;
        pushNode(n, id);
    }
    void pushNode(xml::Node n, int id)
    {
        n.attrs.push_back({"id", std::to_string(id >= 0 ? id : maxId() + 1)});
        m_root.children.push_back(std::move(n));
    }
    static void write(std::ostream &f, const xml::Node &n, int depth)
    {
        const std::string ind(depth, ' ');
        f << ind << "<" << n.name;
        for (auto &a : n.attrs) f << " " << a.first << "=\"" << xml::escape(a.second) << "\"";
        if (n.children.empty() && n.text.empty()) {
            f << "/>\n";
            return;
        }
        f << ">";
        if (!n.children.empty()) {
            f << "\n";
            if (!n.text.empty()) f << xml::escape(n.text) << "\n";
            for (auto &c : n.children) write(f, c, depth + 1);
            f << ind;
        } else {
            f << xml::escape(n.text);
        }
        f << "</" << n.name << ">\n";
    }

    static void formatCheck(const xml::Node &n)
    {
        const std::string fmt = n.attr_or("format", "ascii");
        if (fmt != "ascii" && fmt != "ASCII") throw std::runtime_error("gsFileData: only ASCII payloads are read");
    }

    gsTensorBSplinePatch geometry(const xml::Node &g) const
    {
        const std::string type = g.attr_or("type");
        if (type != "TensorBSpline2" && type != "TensorBSpline3")
            throw std::runtime_error("gsFileData: geometry type '" + type +
                                     "' is not supported (TensorBSpline2/3 only)");
        const int d = type.back() - '0';
        const xml::Node *basis = g.child("Basis");
        if (!basis) throw std::runtime_error("gsFileData: geometry without <Basis>");
        std::vector<gsKnotVector> kv(d);
        std::vector<bool> seen(d, false);
        for (auto &b : basis->children) {
            if (b.name != "Basis") continue;
            if (b.attr_or("type") != "BSplineBasis")
                throw std::runtime_error("gsFileData: only BSplineBasis components are supported");
            const int k = std::atoi(b.attr_or("index", "0").c_str());
            const xml::Node *knots = b.child("KnotVector");
            if (!knots || k < 0 || k >= d) throw std::runtime_error("gsFileData: bad BSplineBasis");
            kv[k] = gsKnotVector(xml::numbers(knots->text), std::atoi(knots->attr_or("degree", "0").c_str()));
            seen[k] = true;
        }
        for (int k = 0; k < d; ++k)
            if (!seen[k]) throw std::runtime_error("gsFileData: missing BSplineBasis component");
        const xml::Node *coefs = g.child("coefs");
        if (!coefs) throw std::runtime_error("gsFileData: geometry without <coefs>");
        formatCheck(*coefs);
        const int gd = std::atoi(coefs->attr_or("geoDim", std::to_string(d)).c_str());
        std::vector<double> v = xml::numbers(coefs->text);
        index_t nb = 1;
        for (auto &k : kv) nb *= k.numBasis();
        if ((index_t)v.size() != nb * gd) throw std::runtime_error("gsFileData: coefficient count mismatch");
        gsMatrix<real_t> C(nb, gd);
        for (index_t i = 0; i < nb; ++i)
            for (int r = 0; r < gd; ++r) C(i, r) = v[(size_t)i * gd + r];
        return gsTensorBSplinePatch(kv, C);
    }

    gsMultiPatch<T> multiPatch(const xml::Node &n) const
    {
        const xml::Node *patches = n.child("patches");
        if (!patches) throw std::runtime_error("gsFileData: MultiPatch without <patches>");
        std::vector<long> ids = xml::integers(patches->text);
        const std::string pt = patches->attr_or("type", "id_range");
        if (pt == "id_range") {
            if (ids.size() != 2) throw std::runtime_error("gsFileData: id_range needs two ids");
            const long a = ids[0], b = ids[1];
            ids.clear();
            for (long k = a; k <= b; ++k) ids.push_back(k);
        } else if (pt != "id_index") {
            throw std::runtime_error("gsFileData: patches type '" + pt + "' not supported");
        }
        std::map<long, int> local;
        gsMultiPatch<T> mp;
        for (size_t k = 0; k < ids.size(); ++k) {
            local[ids[k]] = (int)k;
            mp.addPatch(geometry(byId((int)ids[k])));
        }
        const int d = mp.dim();
        std::vector<typename gsMultiPatch<T>::Interface> ifs;
        for (auto &c : n.children) {
            if (c.name != "interfaces") continue;
            std::vector<long> v = xml::integers(c.text);
            const size_t rec = 4 + 2 * (size_t)d;
            if (v.size() % rec) throw std::runtime_error("gsFileData: malformed <interfaces>");
            for (size_t r = 0; r < v.size(); r += rec) {
                typename gsMultiPatch<T>::Interface f{};
                f.p1 = local.at(v[r]);
                f.s1 = (int)v[r + 1];
                f.p2 = local.at(v[r + 2]);
                f.s2 = (int)v[r + 3];
                const int dir1 = (f.s1 - 1) / 2;
                int t = 0, m[2] = {0, 0}, fl[2] = {0, 0};
                for (int j = 0; j < d; ++j) {
                    if (j == dir1) continue;
                    m[t] = (int)v[r + 4 + j];
                    fl[t] = v[r + 4 + d + j] ? 0 : 1;
                    ++t;
                }
                f.m0 = m[0];
                f.f0 = fl[0];
                f.m1 = m[1];
                f.f1 = fl[1];
                ifs.push_back(f);
            }
        }
        std::vector<std::pair<int, int>> bdr;
        std::map<std::string, std::vector<std::pair<int, int>>> named;
        for (auto &c : n.children) {
            if (c.name != "boundary") continue;
            std::vector<long> v = xml::integers(c.text);
            if (v.size() % 2) throw std::runtime_error("gsFileData: malformed <boundary>");
            for (size_t r = 0; r < v.size(); r += 2) {
                const std::pair<int, int> ps(local.at(v[r]), (int)v[r + 1]);
                bdr.push_back(ps);
                if (c.attr("name")) named[*c.attr("name")].push_back(ps);
            }
        }
        mp.setTopology(ifs, bdr, named);
        return mp;
    }

    gsFunctionExpr<T> function(const xml::Node &n) const
    {
        if (n.attr_or("type") != "FunctionExpr")
            throw std::runtime_error("gsFileData: function type '" + n.attr_or("type") + "' not supported");
        const short_t dim = (short_t)std::atoi(n.attr_or("dim", "2").c_str());
        if (!n.children.empty()) {
            // vector-valued: <c index="i"> components (Neumann fluxes)
            std::vector<std::string> comp;
            for (auto &c : n.children)
                if (c.name == "c") {
                    const size_t i = (size_t)std::atoi(c.attr_or("index", "0").c_str());
                    if (comp.size() <= i) comp.resize(i + 1);
                    comp[i] = c.text;
                }
            return gsFunctionExpr<T>(comp, dim);
        }
        return gsFunctionExpr<T>(n.text, dim);
    }

    gsOptionList options(const xml::Node &n) const
    {
        gsOptionList o;
        for (auto &c : n.children) {
            const std::string label = c.attr_or("label"), value = c.attr_or("value");
            if (c.name == "int") o.addInt(label, c.attr_or("desc"), std::atoi(value.c_str()));
            else if (c.name == "real") o.addReal(label, c.attr_or("desc"), std::atof(value.c_str()));
            else if (c.name == "bool" || c.name == "switch")
                o.addSwitch(label, c.attr_or("desc"), value == "1" || value == "true");
            else if (c.name == "string") o.addString(label, c.attr_or("desc"), value);
        }
        return o;
    }

    gsMatrix<T> matrix(const xml::Node &n) const
    {
        formatCheck(n);
        const index_t r = std::atoi(n.attr_or("rows", "0").c_str()), c = std::atoi(n.attr_or("cols", "0").c_str());
        std::vector<double> v = xml::numbers(n.text);
        if ((index_t)v.size() != r * c) throw std::runtime_error("gsFileData: matrix size mismatch");
        gsMatrix<T> m(r, c);
        for (index_t i = 0; i < r; ++i)
            for (index_t j = 0; j < c; ++j) m(i, j) = v[(size_t)i * c + j];
        return m;
    }

    gsSparseMatrix<T> sparseMatrix(const xml::Node &n) const
    {
        const index_t r = std::atoi(n.attr_or("rows", "0").c_str()), c = std::atoi(n.attr_or("cols", "0").c_str());
        std::vector<double> v = xml::numbers(n.text);
        if (v.size() % 3) throw std::runtime_error("gsFileData: malformed <SparseMatrix>");
        // triplets -> compressed columns; duplicates are summed (setFromTriplets)
        std::vector<std::map<index_t, T>> cols(c);
        for (size_t k = 0; k < v.size(); k += 3) {
            const index_t i = (index_t)v[k], j = (index_t)v[k + 1];
            if (i < 0 || i >= r || j < 0 || j >= c) throw std::runtime_error("gsFileData: entry out of range");
            cols[j][i] += (T)v[k + 2];
        }
        std::vector<index_t> outer(c + 1, 0), inner;
        std::vector<T> val;
        for (index_t j = 0; j < c; ++j) {
            for (auto &e : cols[j]) {
                inner.push_back(e.first);
                val.push_back(e.second);
            }
            outer[j + 1] = (index_t)inner.size();
        }
        gsSparseMatrix<T> m;
        m.setFromCompressed(r, c, std::move(outer), std::move(inner), std::move(val));
        return m;
    }

    gsBoundaryConditions<T> conditions(const xml::Node &n) const
    {
        gsBoundaryConditions<T> bc;
        const std::string mpid = n.attr_or("multipatch");
        gsMultiPatch<T> mp;
        if (!mpid.empty()) mp = multiPatch(byId(std::atoi(mpid.c_str())));
        std::map<int, const xml::Node *> funcs;
        for (auto &c : n.children)
            if (c.name == "Function") funcs[std::atoi(c.attr_or("index", "0").c_str())] = &c;
        for (auto &c : n.children) {
            if (c.name != "bc") continue;
            const std::string type = c.attr_or("type");
            std::vector<std::pair<int, int>> sides;
            if (c.attr("name")) {
                auto it = mp.namedBoundaries().find(*c.attr("name"));
                if (it != mp.namedBoundaries().end()) sides = it->second;
            }
            std::vector<long> v = xml::integers(c.text);
            for (size_t r = 0; r + 1 < v.size(); r += 2) sides.push_back({(int)v[r], (int)v[r + 1]});
            const bool neu = type == "Neumann" || type == "neumann";
            if (!neu && type != "Dirichlet" && type != "dirichlet")
                throw std::runtime_error("gsFileData: boundary condition type '" + type + "' not supported");
            const int fi = std::atoi(c.attr_or("function", "0").c_str());
            gsFunctionExpr<T> g("0", mp.dim() ? mp.dim() : 2);
            if (funcs.count(fi)) g = function(*funcs[fi]);
            for (auto &s : sides)
                bc.addCondition(s.first, (boundary::side)s.second, neu ? condition_type::neumann : condition_type::dirichlet,
                                &g);
        }
        return bc;
    }
};

}  // namespace gismo
