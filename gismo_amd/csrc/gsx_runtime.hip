// Host runtime of libgsx: the C ABI declared in include/gsx.h.
//
// Owns device memory, builds the DOF mapper (host, integer), uploads the
// patches, drives the assembly kernels and the CG / mixed-precision CG
// iterations.  All device work of a context runs on its one HIP stream.
#include "gsx.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <deque>
#include <type_traits>
#include <vector>

#include "gsx_common.h"
#include "gsx_dense.h"
#include "gsx_kernels.h"

using gsx::CGState;
using gsx::GaussTab;
using gsx::PatchDev;

namespace {

using DevPtr = std::shared_ptr<void>;

// Device memory of a context: bytes in use plus a cache of released blocks
// keyed by (rounded) size, so repeated assemblies and solves of the same
// shapes reuse memory instead of calling hipMalloc/hipFree each time.
// Blocks may be released from another host thread than the one that uses
// the context (a matrix assembled on one context and solved on another),
// so the cache is locked.
struct Counter {
    std::atomic<int64_t> bytes{0};     // handed out
    int64_t cached = 0;                // held in the cache
    std::multimap<size_t, void *> cache;
    std::mutex mu;
    ~Counter()
    {
        for (auto &kv : cache) (void)hipFree(kv.second);
    }
    static size_t round(size_t b)
    {
        const size_t g = b >= (1u << 20) ? (size_t)(2u << 20) : 256;
        return (b + g - 1) / g * g;
    }
    void trim()
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto &kv : cache) (void)hipFree(kv.second);
        cache.clear();
        cached = 0;
    }
    void *take(size_t bytes)
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(bytes);
        if (it == cache.end()) return nullptr;
        void *p = it->second;
        cache.erase(it);
        cached -= (int64_t)bytes;
        return p;
    }
    void give(size_t bytes, void *p)
    {
        std::lock_guard<std::mutex> lk(mu);
        cache.emplace(bytes, p);
        cached += (int64_t)bytes;
    }
};

}  // namespace

struct gsx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool timing = false;
    std::string timing_filter;  // empty: time every scope
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<int64_t, double>> stats;
    std::shared_ptr<Counter> mem = std::make_shared<Counter>();
    GaussTab gt;
    int64_t elem_budget = int64_t(16) << 30;   // bytes of element matrices per assembly pass
    int64_t coarse_dense_max = 32768;          // largest coarsest level solved by its dense inverse
    void *scratch = nullptr;   // this device's address of the allocation-free workspace
};

namespace {

gsx_status fail(gsx_ctx *c, gsx_status s, const std::string &msg)
{
    if (c) c->err = msg;
    return s;
}

#define HIP_OK(ctx, expr)                                                                      \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail((ctx), GSX_ERR_HIP,                                                    \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                   \
    } while (0)

#define GSX_TRY(expr)                   \
    do {                                \
        gsx_status _s = (expr);         \
        if (_s != GSX_OK) return _s;    \
    } while (0)

DevPtr dalloc(gsx_ctx *c, size_t bytes)
{
    void *p = nullptr;
    bytes = Counter::round(bytes ? bytes : 16);
    auto mem = c->mem;
    p = mem->take(bytes);
    if (!p && hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        // release the cache and retry once
        mem->trim();
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    mem->bytes += (int64_t)bytes;
    return DevPtr(p, [mem, bytes](void *q) {
        mem->bytes -= (int64_t)bytes;
        mem->give(bytes, q);
    });
}

template <typename T>
T *dp(const DevPtr &p) { return static_cast<T *>(p.get()); }

// Kernel timing scope: events on the context stream when timing is on.
struct KScope {
    gsx_ctx *c;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    bool on = false;
    KScope(gsx_ctx *c_, const char *n) : c(c_), name(n)
    {
        on = c->timing && (c->timing_filter.empty() || c->timing_filter == n);
        if (!on) return;
        a = get();
        b = get();
        (void)hipEventRecord(a, c->stream);
    }
    ~KScope()
    {
        if (!on) return;
        (void)hipEventRecord(b, c->stream);
        c->recs.push_back({name, a, b});
    }
    hipEvent_t get()
    {
        if (!c->pool.empty()) {
            hipEvent_t e = c->pool.back();
            c->pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
};

void resolve_timing(gsx_ctx *c)
{
    for (auto &r : c->recs) {
        (void)hipEventSynchronize(r.b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        auto &s = c->stats[r.name];
        s.first += 1;
        s.second += (double)ms;
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
}

// Gauss-Legendre nodes by Newton's method on P_n in long double, ascending,
// symmetric (the values of the 30-digit tables of gsGaussRule::
// lookupReference, src/gsAssembler/gsGaussRule.hpp:218-547, rounded).
void gauss_rule(int n, double *x, double *w)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        long double z = cosl(pi * ((long double)i + 0.75L) / ((long double)n + 0.5L));
        long double dp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = 0.0L;
            for (int k = 1; k <= n; ++k) {
                const long double p2 = p1;
                p1 = p0;
                p0 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p2) / (long double)k;
            }
            dp = (long double)n * (z * p0 - p1) / (z * z - 1.0L);
            const long double dz = p0 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-30L) break;
        }
        long double p0 = 1.0L, p1 = 0.0L;
        for (int k = 1; k <= n; ++k) {
            const long double p2 = p1;
            p1 = p0;
            p0 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p2) / (long double)k;
        }
        dp = (long double)n * (z * p0 - p1) / (z * z - 1.0L);
        const long double wi = 2.0L / ((1.0L - z * z) * dp * dp);
        if (n % 2 == 1 && i == n / 2) z = 0.0L;
        x[n - 1 - i] = (double)z;
        w[n - 1 - i] = (double)wi;
        x[i] = -(double)z;
        w[i] = (double)wi;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// problem
// ---------------------------------------------------------------------------
struct HostPatch {
    int d;
    int p[3];
    int nb[3];
    int nel[3];
    int64_t nbasis, nelem;
};

struct Group {
    int d, p;
    std::vector<int> patches;
    std::vector<int64_t> prefix;
    int64_t nelem = 0;
    DevPtr d_gpatch, d_gprefix;
};

struct gsx_problem {
    gsx_ctx *ctx = nullptr;
    int d = 0, npatch = 0;
    std::vector<HostPatch> hp;
    std::vector<int64_t> boff;     // basis offset per patch
    std::vector<int> dofs;         // concatenated global index
    int nfree = 0;
    int64_t nbasis = 0, nelem = 0, nglued = 0;
    std::vector<PatchDev> pdev;
    DevPtr d_patches, d_dofs;
    std::vector<DevPtr> keep;      // knots, ctrl, kappa, felem
    DevPtr d_occ_ptr, d_occ_patch, d_occ_local;
    std::vector<Group> groups;
    // pattern cache
    bool have_pattern = false;
    DevPtr d_rowptr, d_col;
    int64_t nnz = 0;
    int maxrow = 0;
    // single patch: free DOF -> index on the tensor grid padded by p zero
    // layers (structured operator of the solvers)
    // host copies of the inputs needed to build coarser levels
    std::vector<std::vector<double>> knots_h;   // npatch * 3 (empty beyond d)
    std::vector<std::vector<double>> ctrl_h;    // control points per patch (nbasis x d)
    std::vector<int> dir_sides_h, ifaces_h;
    DevPtr d_f2g;
    int grid_dims[3] = {1, 1, 1};
    int box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};   // free functions per direction
    // element scratch
    DevPtr d_detmin, d_scan_bsum, d_scan_total;
    int64_t kdoubles = 0, fdoubles = 0;
    double min_det = 0.0;
    // boundary data: values of the eliminated DOFs (global index - nfree;
    // null: homogeneous) and the Neumann load over the free DOFs (null: none)
    DevPtr d_fixed, d_neu;
    // element matrices of the last assembly (GSX_ASM_KEEP_ELEMENTS), laid
    // out at kmat_off, and the serial of that assembly
    DevPtr kept_K;
    uint64_t asm_serial = 0, kept_serial = 0;
    bool zero_kappa = false;        // some element has kappa exactly 0
    // geometry, kappa and source of another problem (a level of a
    // p-multigrid hierarchy): this problem's patches give only the basis
    gsx_problem *geo = nullptr;
};

// DOF numbering with the rules of gsDofMapper::finalizeComp
// (src/gsCore/gsDofMapper.cpp:271-340): standard free DOFs (never matched,
// not eliminated) in (patch, local) scan order, then the coupled free DOFs
// by first appearance, then the eliminated ones by first appearance.  Groups
// of matched DOFs are the connected components of the interface matching
// (gsDofMapper::matchDof/mergeDofsGlobally, gsDofMapper.cpp:99-157,445-478);
// a group is eliminated if any member lies on a Dirichlet side.
static int build_mapper(gsx_problem *P, int ndir, const int32_t *dir_sides, int nif,
                        const int32_t *ifaces, std::string &err)
{
    const int64_t total = P->nbasis;
    std::vector<int64_t> parent(total);
    std::iota(parent.begin(), parent.end(), 0);
    std::vector<char> matched(total, 0), elim(total, 0);
    auto find = [&](int64_t a) {
        while (parent[a] != a) {
            parent[a] = parent[parent[a]];
            a = parent[a];
        }
        return a;
    };
    auto side_list = [&](int k, int side, std::vector<int64_t> &out) {
        const HostPatch &h = P->hp[k];
        const int dir = (side - 1) / 2, par = (side - 1) % 2;
        const int fix = par ? h.nb[dir] - 1 : 0;
        int64_t tot = 1;
        for (int j = 0; j < h.d; ++j)
            if (j != dir) tot *= h.nb[j];
        out.clear();
        for (int64_t t = 0; t < tot; ++t) {
            int idx[3] = {0, 0, 0};
            int64_t r = t;
            for (int j = 0; j < h.d; ++j) {
                if (j == dir) { idx[j] = fix; continue; }
                idx[j] = (int)(r % h.nb[j]);
                r /= h.nb[j];
            }
            out.push_back(idx[0] + (int64_t)h.nb[0] * (idx[1] + (int64_t)h.nb[1] * idx[2]));
        }
    };
    std::vector<int64_t> lst;
    for (int f = 0; f < nif; ++f) {
        const int32_t *I = ifaces + 8 * f;
        const int p1 = I[0], s1 = I[1], p2 = I[2], s2 = I[3];
        if (p1 < 0 || p1 >= P->npatch || p2 < 0 || p2 >= P->npatch || s1 < 1 || s1 > 2 * P->d ||
            s2 < 1 || s2 > 2 * P->d) {
            err = "invalid interface record";
            return -1;
        }
        const HostPatch &A = P->hp[p1], &B = P->hp[p2];
        const int dir1 = (s1 - 1) / 2, dir2 = (s2 - 1) / 2, par2 = (s2 - 1) % 2;
        side_list(p1, s1, lst);
        for (int64_t la : lst) {
            int ia[3] = {(int)(la % A.nb[0]), (int)((la / A.nb[0]) % A.nb[1]),
                         A.d == 3 ? (int)(la / ((int64_t)A.nb[0] * A.nb[1])) : 0};
            int ib[3] = {0, 0, 0};
            ib[dir2] = par2 ? B.nb[dir2] - 1 : 0;
            int j = 0;
            for (int k = 0; k < A.d; ++k) {
                if (k == dir1) continue;
                const int m = I[4 + 2 * j], flip = I[5 + 2 * j];
                if (m < 0 || m >= B.d || A.nb[k] != B.nb[m]) {
                    err = "non-conforming interface";
                    return -1;
                }
                ib[m] = flip ? (B.nb[m] - 1 - ia[k]) : ia[k];
                ++j;
            }
            const int64_t lb = ib[0] + (int64_t)B.nb[0] * (ib[1] + (int64_t)B.nb[1] * ib[2]);
            const int64_t ga = P->boff[p1] + la, gb = P->boff[p2] + lb;
            matched[ga] = matched[gb] = 1;
            const int64_t ra = find(ga), rb = find(gb);
            if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
        }
    }
    for (int s = 0; s < ndir; ++s) {
        const int k = dir_sides[2 * s], side = dir_sides[2 * s + 1];
        if (k < 0 || k >= P->npatch || side < 1 || side > 2 * P->d) {
            err = "invalid Dirichlet side";
            return -1;
        }
        side_list(k, side, lst);
        for (int64_t l : lst) elim[P->boff[k] + l] = 1;
    }
    std::vector<char> comp_elim(total, 0);
    for (int64_t i = 0; i < total; ++i)
        if (elim[i]) comp_elim[find(i)] = 1;
    int64_t nstd = 0, ncpl = 0;
    std::vector<int64_t> comp_id(total, -1);
    for (int64_t i = 0; i < total; ++i) {
        const int64_t r = find(i);
        if (comp_elim[r]) continue;
        if (!matched[i]) ++nstd;
        else if (comp_id[r] < 0) { comp_id[r] = 0; ++ncpl; }
    }
    std::fill(comp_id.begin(), comp_id.end(), -1);
    const int64_t nfree = nstd + ncpl;
    if (nfree >= (int64_t)INT32_MAX) {
        err = "too many DOFs for int32 indices";
        return -1;
    }
    int64_t cur_std = 0, cur_cpl = nstd, cur_elim = nfree;
    P->dofs.assign(total, -1);
    for (int64_t i = 0; i < total; ++i) {
        const int64_t r = find(i);
        if (comp_elim[r]) {
            if (comp_id[r] < 0) comp_id[r] = cur_elim++;
            P->dofs[i] = (int)comp_id[r];
        } else if (!matched[i]) {
            P->dofs[i] = (int)cur_std++;
        } else {
            if (comp_id[r] < 0) comp_id[r] = cur_cpl++;
            P->dofs[i] = (int)comp_id[r];
        }
    }
    P->nglued = cur_elim;
    return (int)nfree;
}

extern "C" {

gsx_status gsx_gauss_legendre(int32_t n, double *x, double *w)
{
    if (n < 1 || n > 64 || !x || !w) return GSX_ERR_ARG;
    gauss_rule(n, x, w);
    return GSX_OK;
}

gsx_status gsx_mapper_build(int32_t d, int32_t npatch, const int32_t *nbasis_dir, int32_t ndir,
                            const int32_t *dir_sides, int32_t nif, const int32_t *ifaces,
                            int32_t *dofs_out, int64_t *nfree, int64_t *nglued)
{
    if ((d != 2 && d != 3) || npatch <= 0 || !nbasis_dir || !dofs_out || (ndir > 0 && !dir_sides) ||
        (nif > 0 && !ifaces))
        return GSX_ERR_ARG;
    gsx_problem P;
    P.d = d;
    P.npatch = npatch;
    int64_t off = 0;
    for (int k = 0; k < npatch; ++k) {
        HostPatch h{};
        h.d = d;
        h.nbasis = 1;
        for (int j = 0; j < 3; ++j) {
            h.nb[j] = j < d ? nbasis_dir[3 * k + j] : 1;
            if (h.nb[j] < 1) return GSX_ERR_ARG;
            h.nbasis *= h.nb[j];
        }
        P.hp.push_back(h);
        P.boff.push_back(off);
        off += h.nbasis;
    }
    P.nbasis = off;
    std::string err;
    const int nf = build_mapper(&P, ndir, dir_sides, nif, ifaces, err);
    if (nf < 0) return GSX_ERR_ARG;
    std::memcpy(dofs_out, P.dofs.data(), sizeof(int) * off);
    if (nfree) *nfree = nf;
    if (nglued) *nglued = P.nglued;
    return GSX_OK;
}

// Workspace of gsx_cg_solve_noalloc: a fixed device array, one copy per GPU,
// alive until the process ends.  Host code uses only the address that
// hipGetSymbolAddress returns for the current device (kept per device in
// g_scratch_addr and in every context of that device), never the symbol.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static void *g_scratch_addr[64] = {};

// device address of `scratch` on the context's device (looked up once per
// device, with that device current)
static gsx_status scratch_setup(gsx_ctx *c)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device < 0 || c->device >= 64) return fail(c, GSX_ERR_ARG, "device index out of range");
    if (!g_scratch_addr[c->device]) {
        void *ptr = nullptr;
        HIP_OK(c, hipSetDevice(c->device));
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(c, GSX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[c->device] = ptr;
    }
    c->scratch = g_scratch_addr[c->device];
    return GSX_OK;
}

gsx_status gsx_ctx_create(int32_t device, gsx_ctx **out)
{
    if (!out) return GSX_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return GSX_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) return GSX_ERR_ARG;
    gsx_ctx *c = new gsx_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return GSX_ERR_HIP;
    }
    std::memset(&c->gt, 0, sizeof(c->gt));
    for (int q = 1; q <= gsx::kMaxQ1; ++q) gauss_rule(q, c->gt.x[q], c->gt.w[q]);
    if (scratch_setup(c) != GSX_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return GSX_ERR_HIP;
    }
    *out = c;
    return GSX_OK;
}

gsx_status gsx_ctx_set_elem_budget(gsx_ctx *c, int64_t bytes)
{
    if (!c || bytes <= 0) return GSX_ERR_ARG;
    c->elem_budget = bytes;
    return GSX_OK;
}

gsx_status gsx_ctx_set_coarse_dense_max(gsx_ctx *c, int64_t rows)
{
    if (!c || rows < 1 || rows > 32768) return GSX_ERR_ARG;
    c->coarse_dense_max = rows;
    return GSX_OK;
}

gsx_status gsx_ctx_arch(const gsx_ctx *c, char *buf, int32_t cap)
{
    if (!c || !buf || cap < 1) return GSX_ERR_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) return GSX_ERR_HIP;
    std::snprintf(buf, (size_t)cap, "%s", prop.gcnArchName);
    return GSX_OK;
}

gsx_status gsx_ctx_destroy(gsx_ctx *c)
{
    if (!c) return GSX_ERR_ARG;
    (void)hipStreamSynchronize(c->stream);
    for (auto &r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return GSX_OK;
}

const char *gsx_last_error(const gsx_ctx *c) { return c ? c->err.c_str() : "null context"; }

gsx_status gsx_ctx_synchronize(gsx_ctx *c)
{
    if (!c) return GSX_ERR_ARG;
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return GSX_OK;
}

gsx_status gsx_ctx_enable_timing(gsx_ctx *c, int32_t enable)
{
    if (!c) return GSX_ERR_ARG;
    c->timing = enable != 0;
    return GSX_OK;
}

gsx_status gsx_ctx_timing_filter(gsx_ctx *c, const char *name)
{
    if (!c) return GSX_ERR_ARG;
    c->timing_filter = name ? name : "";
    return GSX_OK;
}

gsx_status gsx_ctx_reset_timing(gsx_ctx *c)
{
    if (!c) return GSX_ERR_ARG;
    resolve_timing(c);
    c->stats.clear();
    return GSX_OK;
}

gsx_status gsx_ctx_kernel_stats(gsx_ctx *c, const char *name, int64_t *launches, double *total_ms)
{
    if (!c || !name) return GSX_ERR_ARG;
    resolve_timing(c);
    auto it = c->stats.find(name);
    if (launches) *launches = it == c->stats.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.second;
    return GSX_OK;
}

gsx_status gsx_device_alloc(gsx_ctx *c, int64_t bytes, void **out)
{
    if (!c || !out || bytes < 0) return GSX_ERR_ARG;
    HIP_OK(c, hipMalloc(out, (size_t)(bytes ? bytes : 16)));
    c->mem->bytes += bytes ? bytes : 16;
    return GSX_OK;
}

gsx_status gsx_device_free(gsx_ctx *c, void *p)
{
    if (!c) return GSX_ERR_ARG;
    if (!p) return GSX_OK;
    size_t sz = 0;
    (void)hipMemPtrGetInfo(p, &sz);
    HIP_OK(c, hipFree(p));
    c->mem->bytes -= (int64_t)sz;
    return GSX_OK;
}

gsx_status gsx_copy_to_device(gsx_ctx *c, void *dev, const void *host, int64_t bytes)
{
    if (!c || (bytes && (!dev || !host))) return GSX_ERR_ARG;
    HIP_OK(c, hipMemcpyAsync(dev, host, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return GSX_OK;
}

gsx_status gsx_copy_to_host(gsx_ctx *c, void *host, const void *dev, int64_t bytes)
{
    if (!c || (bytes && (!dev || !host))) return GSX_ERR_ARG;
    HIP_OK(c, hipMemcpyAsync(host, dev, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return GSX_OK;
}

gsx_status gsx_ctx_memory(gsx_ctx *c, int64_t *bytes)
{
    if (!c || !bytes) return GSX_ERR_ARG;
    *bytes = c->mem->bytes;
    return GSX_OK;
}

// ---------------------------------------------------------------------------
gsx_status gsx_problem_create(gsx_ctx *c, int32_t npatch, const gsx_patch_desc *patches,
                              int32_t ndir, const int32_t *dir_sides, int32_t nif,
                              const int32_t *ifaces, gsx_problem **out)
{
    if (!c || !patches || npatch <= 0 || !out || (ndir > 0 && !dir_sides) || (nif > 0 && !ifaces))
        return fail(c, GSX_ERR_ARG, "gsx_problem_create: null argument");
    auto P = std::make_unique<gsx_problem>();
    P->ctx = c;
    P->npatch = npatch;
    P->d = patches[0].d;
    if (P->d != 2 && P->d != 3) return fail(c, GSX_ERR_ARG, "dimension must be 2 or 3");
    int64_t boff = 0, eoff = 0;
    for (int k = 0; k < npatch; ++k) {
        const gsx_patch_desc &D = patches[k];
        if (D.d != P->d) return fail(c, GSX_ERR_ARG, "all patches must share the dimension");
        HostPatch h{};
        h.d = D.d;
        h.nbasis = 1;
        h.nelem = 1;
        for (int j = 0; j < D.d; ++j) {
            const int p = D.p[j], nk = D.nknots[j];
            const double *U = D.knots_host[j];
            if (p < 1 || p > gsx::kMaxDeg) return fail(c, GSX_ERR_ARG, "degree must be in 1..4");
            if (p != D.p[0]) return fail(c, GSX_ERR_ARG, "degree must be equal in all directions");
            if (!U || nk < 2 * p + 2) return fail(c, GSX_ERR_ARG, "knot vector too short");
            for (int i = 0; i < p; ++i)
                if (U[i] != U[p] || U[nk - 1 - i] != U[nk - 1 - p])
                    return fail(c, GSX_ERR_ARG, "knot vectors must be open (clamped)");
            for (int i = p; i < nk - p - 1; ++i)
                if (!(U[i + 1] > U[i]))
                    return fail(c, GSX_ERR_ARG, "interior knots must be strictly increasing");
            h.p[j] = p;
            h.nb[j] = nk - p - 1;
            h.nel[j] = nk - 2 * p - 1;
            h.nbasis *= h.nb[j];
            h.nelem *= h.nel[j];
        }
        for (int j = D.d; j < 3; ++j) { h.p[j] = 0; h.nb[j] = 1; h.nel[j] = 1; }
        if (!D.ctrl_host) return fail(c, GSX_ERR_ARG, "missing control points");
        if (D.fkind == 0 && !D.felem_host) return fail(c, GSX_ERR_ARG, "fkind 0 needs felem");
        P->hp.push_back(h);
        P->boff.push_back(boff);
        boff += h.nbasis;
        eoff += h.nelem;
    }
    P->nbasis = boff;
    P->knots_h.resize((size_t)npatch * 3);
    for (int k = 0; k < npatch; ++k)
        for (int j = 0; j < patches[k].d; ++j)
            P->knots_h[3 * k + j].assign(patches[k].knots_host[j], patches[k].knots_host[j] + patches[k].nknots[j]);
    P->ctrl_h.resize(npatch);
    for (int k = 0; k < npatch; ++k)
        P->ctrl_h[k].assign(patches[k].ctrl_host, patches[k].ctrl_host + P->hp[k].nbasis * patches[k].d);
    if (ndir > 0) P->dir_sides_h.assign(dir_sides, dir_sides + 2 * ndir);
    if (nif > 0) P->ifaces_h.assign(ifaces, ifaces + 8 * nif);
    P->nelem = eoff;
    std::string err;
    const int nfree = build_mapper(P.get(), ndir, dir_sides, nif, ifaces, err);
    if (nfree < 0) return fail(c, GSX_ERR_ARG, err);
    P->nfree = nfree;

    // upload
    hipStream_t s = c->stream;
    P->d_dofs = dalloc(c, sizeof(int) * P->nbasis);
    if (!P->d_dofs) return fail(c, GSX_ERR_OOM, "device allocation (dofs)");
    HIP_OK(c, hipMemcpyAsync(P->d_dofs.get(), P->dofs.data(), sizeof(int) * P->nbasis,
                             hipMemcpyHostToDevice, s));
    eoff = 0;
    int64_t kdoubles = 0, fdoubles = 0;
    for (int k = 0; k < npatch; ++k) {
        const gsx_patch_desc &D = patches[k];
        const HostPatch &h = P->hp[k];
        PatchDev pd{};
        pd.d = h.d;
        for (int j = 0; j < 3; ++j) {
            pd.p[j] = h.p[j];
            pd.nb[j] = h.nb[j];
            pd.nel[j] = h.nel[j];
            pd.knots[j] = nullptr;
        }
        for (int j = 0; j < h.d; ++j) {
            DevPtr kb = dalloc(c, sizeof(double) * D.nknots[j]);
            if (!kb) return fail(c, GSX_ERR_OOM, "device allocation (knots)");
            HIP_OK(c, hipMemcpyAsync(kb.get(), D.knots_host[j], sizeof(double) * D.nknots[j],
                                     hipMemcpyHostToDevice, s));
            pd.knots[j] = dp<double>(kb);
            P->keep.push_back(kb);
        }
        DevPtr cb = dalloc(c, sizeof(double) * h.nbasis * h.d);
        if (!cb) return fail(c, GSX_ERR_OOM, "device allocation (ctrl)");
        HIP_OK(c, hipMemcpyAsync(cb.get(), D.ctrl_host, sizeof(double) * h.nbasis * h.d,
                                 hipMemcpyHostToDevice, s));
        pd.ctrl = dp<double>(cb);
        P->keep.push_back(cb);
        pd.kappa = nullptr;
        pd.felem = nullptr;
        if (D.kappa_host) {
            for (int64_t q = 0; q < h.nelem; ++q) P->zero_kappa = P->zero_kappa || D.kappa_host[q] == 0.0;
            DevPtr b = dalloc(c, sizeof(double) * h.nelem);
            if (!b) return fail(c, GSX_ERR_OOM, "device allocation (kappa)");
            HIP_OK(c, hipMemcpyAsync(b.get(), D.kappa_host, sizeof(double) * h.nelem,
                                     hipMemcpyHostToDevice, s));
            pd.kappa = dp<double>(b);
            P->keep.push_back(b);
        }
        if (D.felem_host && D.fkind == 0) {
            DevPtr b = dalloc(c, sizeof(double) * h.nelem);
            if (!b) return fail(c, GSX_ERR_OOM, "device allocation (felem)");
            HIP_OK(c, hipMemcpyAsync(b.get(), D.felem_host, sizeof(double) * h.nelem,
                                     hipMemcpyHostToDevice, s));
            pd.felem = dp<double>(b);
            P->keep.push_back(b);
        }
        pd.fkind = D.fkind;
        // orientation from the control net at the parametric origin: the
        // Jacobian there has columns along the net's first edges (open knots)
        {
            const double *c0 = D.ctrl_host;
            const int d = h.d;
            const int64_t s1 = 1, s2 = h.nb[0], s3 = (int64_t)h.nb[0] * h.nb[1];
            double e[3][3] = {{0}};
            for (int r = 0; r < d; ++r) {
                e[0][r] = c0[s1 * d + r] - c0[r];
                e[1][r] = c0[s2 * d + r] - c0[r];
                if (d == 3) e[2][r] = c0[s3 * d + r] - c0[r];
            }
            const double det = d == 2 ? e[0][0] * e[1][1] - e[0][1] * e[1][0]
                                      : e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                                            e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                                            e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
            pd.orient = det < 0.0 ? -1 : 1;
        }
        pd.dofs = dp<int>(P->d_dofs) + P->boff[k];
        pd.nelem = h.nelem;
        pd.elem_off = eoff;
        int nb1 = h.p[0] + 1;
        int NB = (h.d == 2) ? nb1 * nb1 : nb1 * nb1 * nb1;
        pd.kmat_off = kdoubles;
        pd.fvec_off = fdoubles;
        kdoubles += h.nelem * (int64_t)NB * NB;
        fdoubles += h.nelem * (int64_t)NB;
        eoff += h.nelem;
        P->pdev.push_back(pd);
    }
    P->kdoubles = kdoubles;
    P->fdoubles = fdoubles;
    P->d_patches = dalloc(c, sizeof(PatchDev) * npatch);
    if (!P->d_patches) return fail(c, GSX_ERR_OOM, "device allocation (patches)");
    HIP_OK(c, hipMemcpyAsync(P->d_patches.get(), P->pdev.data(), sizeof(PatchDev) * npatch,
                             hipMemcpyHostToDevice, s));

    // groups by (d, p)
    for (int k = 0; k < npatch; ++k) {
        const HostPatch &h = P->hp[k];
        Group *g = nullptr;
        for (auto &gg : P->groups)
            if (gg.d == h.d && gg.p == h.p[0]) g = &gg;
        if (!g) {
            P->groups.push_back(Group());
            g = &P->groups.back();
            g->d = h.d;
            g->p = h.p[0];
        }
        g->patches.push_back(k);
    }
    for (auto &g : P->groups) {
        g.prefix.assign(1, 0);
        for (int k : g.patches) g.prefix.push_back(g.prefix.back() + P->hp[k].nelem);
        g.nelem = g.prefix.back();
        g.d_gpatch = dalloc(c, sizeof(int) * g.patches.size());
        g.d_gprefix = dalloc(c, sizeof(int64_t) * g.prefix.size());
        if (!g.d_gpatch || !g.d_gprefix) return fail(c, GSX_ERR_OOM, "device allocation (groups)");
        HIP_OK(c, hipMemcpyAsync(g.d_gpatch.get(), g.patches.data(), sizeof(int) * g.patches.size(),
                                 hipMemcpyHostToDevice, s));
        HIP_OK(c, hipMemcpyAsync(g.d_gprefix.get(), g.prefix.data(), sizeof(int64_t) * g.prefix.size(),
                                 hipMemcpyHostToDevice, s));
    }

    // occurrences of every free DOF (patch order, then local order)
    std::vector<int> occ_cnt(nfree + 1, 0);
    for (int64_t i = 0; i < P->nbasis; ++i)
        if (P->dofs[i] < nfree) occ_cnt[P->dofs[i] + 1]++;
    for (int r = 0; r < nfree; ++r) occ_cnt[r + 1] += occ_cnt[r];
    std::vector<int> occ_patch(occ_cnt[nfree]);
    std::vector<int64_t> occ_local(occ_cnt[nfree]);
    std::vector<int> fillp(occ_cnt.begin(), occ_cnt.end() - 1);
    for (int k = 0; k < npatch; ++k)
        for (int64_t l = 0; l < P->hp[k].nbasis; ++l) {
            const int g = P->dofs[P->boff[k] + l];
            if (g < nfree) {
                const int o = fillp[g]++;
                occ_patch[o] = k;
                occ_local[o] = l;
            }
        }
    P->d_occ_ptr = dalloc(c, sizeof(int) * (nfree + 1));
    P->d_occ_patch = dalloc(c, sizeof(int) * std::max<size_t>(1, occ_patch.size()));
    P->d_occ_local = dalloc(c, sizeof(int64_t) * std::max<size_t>(1, occ_local.size()));
    if (!P->d_occ_ptr || !P->d_occ_patch || !P->d_occ_local)
        return fail(c, GSX_ERR_OOM, "device allocation (occurrences)");
    HIP_OK(c, hipMemcpyAsync(P->d_occ_ptr.get(), occ_cnt.data(), sizeof(int) * (nfree + 1),
                             hipMemcpyHostToDevice, s));
    if (!occ_patch.empty()) {
        HIP_OK(c, hipMemcpyAsync(P->d_occ_patch.get(), occ_patch.data(), sizeof(int) * occ_patch.size(),
                                 hipMemcpyHostToDevice, s));
        HIP_OK(c, hipMemcpyAsync(P->d_occ_local.get(), occ_local.data(),
                                 sizeof(int64_t) * occ_local.size(), hipMemcpyHostToDevice, s));
    }
    if (npatch == 1 && nfree > 0) {
        const HostPatch &h = P->hp[0];
        const int p = h.p[0];
        P->grid_dims[0] = h.nb[0] + 2 * p;
        P->grid_dims[1] = h.nb[1] + 2 * p;
        P->grid_dims[2] = h.d == 3 ? h.nb[2] + 2 * p : 1;
        std::vector<int> f2g(nfree, 0);
        for (int64_t l = 0; l < h.nbasis; ++l) {
            const int g = P->dofs[l];
            if (g < nfree) {
                const int ix = (int)(l % h.nb[0]), iy = (int)((l / h.nb[0]) % h.nb[1]);
                const int iz = h.d == 3 ? (int)(l / ((int64_t)h.nb[0] * h.nb[1])) : 0;
                const int pz = h.d == 3 ? p : 0;
                f2g[g] = (ix + p) + P->grid_dims[0] * ((iy + p) + P->grid_dims[1] * (iz + pz));
            }
        }
        // free functions of a single patch form a box: [1 or 0, nb-2 or nb-1]
        int64_t boxsize = 1;
        for (int k = 0; k < 3; ++k) {
            P->box_lo[k] = 0;
            P->box_hi[k] = k < h.d ? h.nb[k] - 1 : 0;
        }
        for (int sidx = 0; sidx < ndir; ++sidx) {
            const int side = dir_sides[2 * sidx + 1], dir = (side - 1) / 2;
            if (side % 2 == 1) P->box_lo[dir] = 1; else P->box_hi[dir] = h.nb[dir] - 2;
        }
        for (int k = 0; k < h.d; ++k) boxsize *= std::max(0, P->box_hi[k] - P->box_lo[k] + 1);
        if (boxsize != nfree) goto no_structure;
        P->d_f2g = dalloc(c, sizeof(int) * nfree);
        if (!P->d_f2g) return fail(c, GSX_ERR_OOM, "device allocation (grid map)");
        HIP_OK(c, hipMemcpyAsync(P->d_f2g.get(), f2g.data(), sizeof(int) * nfree, hipMemcpyHostToDevice, s));
    }
no_structure:
    HIP_OK(c, hipStreamSynchronize(s));
    *out = P.release();
    return GSX_OK;
}

// A problem discretised with the bases in `patches` (degrees and knot
// vectors; control points, kappa and sources are not read) over the geometry
// map, kappa and source of `geo` (gsCDRAssembler(mp, bases[i], ...) of
// examples/pMultiGrid_example.cpp:445-460, where the level bases differ from
// the geometry's).  Each patch's parameter box must be the geometry patch's.
gsx_status gsx_problem_create_geo(gsx_problem *geo, int32_t npatch, const gsx_patch_desc *patches, int32_t ndir,
                                  const int32_t *dir_sides, int32_t nif, const int32_t *ifaces, gsx_problem **out)
{
    if (!geo || !patches || !out) return GSX_ERR_ARG;
    gsx_ctx *c = geo->ctx;
    if (geo->geo) return fail(c, GSX_ERR_ARG, "the geometry problem must carry its own geometry");
    if (npatch != geo->npatch) return fail(c, GSX_ERR_ARG, "patch count differs from the geometry's");
    std::vector<gsx_patch_desc> desc(patches, patches + npatch);
    std::vector<std::vector<double>> zeros((size_t)npatch);
    for (int k = 0; k < npatch; ++k) {
        gsx_patch_desc &D = desc[(size_t)k];
        if (D.d != geo->d) return fail(c, GSX_ERR_ARG, "dimension differs from the geometry's");
        int64_t nbas = 1;
        for (int j = 0; j < D.d; ++j) {
            const std::vector<double> &G = geo->knots_h[(size_t)3 * k + j];
            const int nk = D.nknots[j];
            if (!D.knots_host[j] || nk < 2) return fail(c, GSX_ERR_ARG, "knot vector too short");
            if (D.knots_host[j][0] != G.front() || D.knots_host[j][nk - 1] != G.back())
                return fail(c, GSX_ERR_ARG, "parameter box differs from the geometry patch's");
            nbas *= nk - D.p[j] - 1;
        }
        zeros[(size_t)k].assign((size_t)std::max<int64_t>(1, nbas) * D.d, 0.0);
        D.ctrl_host = zeros[(size_t)k].data();
        D.kappa_host = nullptr;
        D.felem_host = nullptr;
        D.fkind = 1;
    }
    gsx_problem *P = nullptr;
    GSX_TRY(gsx_problem_create(c, npatch, desc.data(), ndir, dir_sides, nif, ifaces, &P));
    P->geo = geo;
    P->zero_kappa = geo->zero_kappa;
    *out = P;
    return GSX_OK;
}

gsx_status gsx_problem_destroy(gsx_problem *P)
{
    if (!P) return GSX_ERR_ARG;
    (void)hipStreamSynchronize(P->ctx->stream);
    delete P;
    return GSX_OK;
}

gsx_status gsx_problem_info(const gsx_problem *P, int64_t *nbasis, int64_t *nfree, int64_t *nelem,
                            int64_t *nglued)
{
    if (!P) return GSX_ERR_ARG;
    if (nbasis) *nbasis = P->nbasis;
    if (nfree) *nfree = P->nfree;
    if (nelem) *nelem = P->nelem;
    if (nglued) *nglued = P->nglued;
    return GSX_OK;
}

gsx_status gsx_problem_dofs(const gsx_problem *P, int32_t *dofs)
{
    if (!P || !dofs) return GSX_ERR_ARG;
    std::memcpy(dofs, P->dofs.data(), sizeof(int) * P->nbasis);
    return GSX_OK;
}

gsx_status gsx_problem_min_detj(const gsx_problem *P, double *m)
{
    if (!P || !m) return GSX_ERR_ARG;
    *m = P->min_det;
    return GSX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// matrix + solver workspace
// ---------------------------------------------------------------------------
struct CGWork {
    int n = 0;
    int nparts = 0;
    DevPtr r, z, p, t;           // FP64 vectors
    DevPtr parts;                // 4 partial arrays
    DevPtr state, hist;
    int hist_cap = 0;
    // mixed
    DevPtr r32, z32, p32, t32, xc32, diag32;
};

// Structured operator of a single-patch matrix (see gsx_stencil.hip):
// coefficients [slot][free row], vectors on the tensor grid padded by p
// zero layers on every side.
struct StencilOp {
    bool available = false;   // single patch: the structure applies
    bool built = false;       // grid workspace allocated
    bool s64_ready = false;   // slots filled (by the gather or by conversion)
    int d = 0, p = 0;
    int nx = 1, ny = 1, nz = 1;   // padded grid
    int m[3] = {1, 1, 1};         // free box (rows lexicographic over it)
    int64_t ngrid = 0;
    int slots = 0;                // s64: all slots; s32: stencil_half(slots)
    std::vector<int> f2g_host;
    DevPtr f2g, s64, s32;
    bool s32_ready = false;
    DevPtr s32full;                   // all slots in FP32 (batch members that ask for it)
    // grid-layout workspace
    DevPtr diag64, diag32;
    DevPtr xg, rg, bg, zg, pg, tg;            // FP64
    DevPtr r32, z32, p32, t32, xc32;          // FP32
    DevPtr parts;
};

struct gsx_matrix {
    gsx_ctx *ctx = nullptr;
    int n = 0;
    int64_t nnz = 0;
    int lpr = 8;
    DevPtr rowptr, col, val, val32, diag;
    CGWork ws;
    StencilOp st;
    bool use_stencil = true;
    bool mixed = false;       // assembled with GSX_MIXED (FP32 operator available)
    const gsx_problem *src = nullptr;   // problem it was assembled from (values untouched since)
    uint64_t src_serial = 0;            // that assembly's serial in the problem
    std::vector<double> hist;
};

static gsx_status pattern_build(gsx_problem *P)
{
    gsx_ctx *c = P->ctx;
    hipStream_t s = c->stream;
    const int n = P->nfree;
    DevPtr counts = dalloc(c, sizeof(int) * (n + 1));
    DevPtr errf = dalloc(c, sizeof(int));
    P->d_rowptr = dalloc(c, sizeof(int) * (n + 1));
    const int nblk = (n + 1023) / 1024 + 1;
    P->d_scan_bsum = dalloc(c, sizeof(long long) * nblk);
    P->d_scan_total = dalloc(c, sizeof(long long));
    if (!counts || !errf || !P->d_rowptr || !P->d_scan_bsum || !P->d_scan_total)
        return fail(c, GSX_ERR_OOM, "device allocation (pattern)");
    HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int), s));
    {
        KScope k(c, "pattern_count");
        HIP_OK(c, gsx::launch_rowcount(dp<PatchDev>(P->d_patches), dp<int>(P->d_occ_ptr),
                                       dp<int>(P->d_occ_patch), dp<int64_t>(P->d_occ_local), n,
                                       dp<int>(counts), dp<int>(errf), s));
    }
    {
        KScope k(c, "pattern_scan");
        HIP_OK(c, gsx::launch_scan(dp<int>(counts), n, dp<int>(P->d_rowptr),
                                   dp<long long>(P->d_scan_bsum), dp<long long>(P->d_scan_total), s));
    }
    int errh = 0;
    long long total = 0;
    HIP_OK(c, hipMemcpyAsync(&errh, errf.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipMemcpyAsync(&total, P->d_scan_total.get(), sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    if (errh) return fail(c, GSX_ERR_ARG, "row of an interface DOF exceeds the pattern buffer");
    if (total >= (long long)INT32_MAX) return fail(c, GSX_ERR_ARG, "nnz exceeds int32 indexing");
    // max row length
    std::vector<int> cnt(n);
    HIP_OK(c, hipMemcpyAsync(cnt.data(), counts.get(), sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    P->maxrow = n ? *std::max_element(cnt.begin(), cnt.end()) : 0;
    P->nnz = total;
    P->d_col = dalloc(c, sizeof(int) * std::max<long long>(1, total));
    if (!P->d_col) return fail(c, GSX_ERR_OOM, "device allocation (col)");
    {
        KScope k(c, "pattern_fill");
        if (P->npatch == 1 && P->d_f2g) {
            const HostPatch &h = P->hp[0];
            HIP_OK(c, gsx::launch_rowfill_patch(h.d, h.p[0], P->box_lo, P->box_hi, h.nb, h.nel,
                                                P->grid_dims, n, dp<int>(P->d_f2g), dp<int>(P->d_rowptr),
                                                dp<int>(P->d_col), s));
        } else {
            HIP_OK(c, gsx::launch_rowfill(dp<PatchDev>(P->d_patches), dp<int>(P->d_occ_ptr),
                                          dp<int>(P->d_occ_patch), dp<int64_t>(P->d_occ_local), n,
                                          dp<int>(P->d_rowptr), dp<int>(P->d_col), s));
        }
    }
    P->have_pattern = true;
    return GSX_OK;
}

extern "C" {

gsx_status gsx_assemble(gsx_problem *P, int32_t mode, int32_t precision, gsx_matrix **A_out,
                        double *rhs_dev)
{
    if (!P || !A_out) return GSX_ERR_ARG;
    gsx_ctx *c = P->ctx;
    const int32_t flags =
        GSX_ASM_REBUILD_PATTERN | GSX_ASM_CSR_ONLY | GSX_ASM_KEEP_ELEMENTS | GSX_ASM_ELEMENTWISE | GSX_ASM_ROWWISE;
    if ((mode & ~flags) != GSX_PATTERN_EXPR && (mode & ~flags) != GSX_PATTERN_FULL)
        return fail(c, GSX_ERR_ARG, "unknown pattern mode");
    if (precision != GSX_FP64 && precision != GSX_MIXED)
        return fail(c, GSX_ERR_ARG, "unknown precision");
    hipStream_t s = c->stream;
    const bool rebuild = (mode & GSX_ASM_REBUILD_PATTERN) != 0;
    const bool csr_only = (mode & GSX_ASM_CSR_ONLY) != 0;
    const bool keep_elements = (mode & GSX_ASM_KEEP_ELEMENTS) != 0;
    const bool rowwise_forced = (mode & GSX_ASM_ROWWISE) != 0;
    bool elementwise = (mode & GSX_ASM_ELEMENTWISE) != 0;
    mode &= ~(GSX_ASM_REBUILD_PATTERN | GSX_ASM_CSR_ONLY | GSX_ASM_KEEP_ELEMENTS | GSX_ASM_ELEMENTWISE |
              GSX_ASM_ROWWISE);
    P->kept_K.reset();
    const uint64_t serial = ++P->asm_serial;
    if (!P->have_pattern || rebuild) {
        gsx_status st = pattern_build(P);
        if (st != GSX_OK) return st;
    }
    if (gsx::gather_maxrow_bucket(P->maxrow) < 0)
        return fail(c, GSX_ERR_ARG, "rows longer than 1024 entries are not supported");
    const bool single = P->npatch == 1 && P->d_f2g && P->nfree > 0;
    if (!P->d_detmin) {
        P->d_detmin = dalloc(c, sizeof(double) * (std::max<int64_t>(1, P->nelem) + 260));
        if (!P->d_detmin) return fail(c, GSX_ERR_OOM, "device allocation (Jacobian check)");
    }
    // single 3D patch: sum-factorised rows, no element matrices
    // (by default where its matrix-core variant runs: degrees 2 to 4)
    if (!rowwise_forced && !gsx::sf_mfma(P->hp[0].p[0])) elementwise = true;
    if (P->geo) {
        // geometry in another problem's basis: element matrices by the
        // two-basis kernel (the row assembly evaluates the map in the
        // discretisation basis)
        elementwise = true;
        if (P->geo->npatch != P->npatch) return fail(c, GSX_ERR_ARG, "geometry problem changed");
    }
    const bool use_sf = single && !elementwise && !keep_elements && !P->zero_kappa && P->hp[0].d == 3 &&
                        P->hp[0].p[0] == P->hp[0].p[1] && P->hp[0].p[0] == P->hp[0].p[2] &&
                        gsx::sf_supported(3, P->hp[0].p[0]);
    // Element matrices live only during the assembly, in a buffer of at most
    // the context's budget: a single patch is assembled in slabs of element
    // layers along its slowest direction (each slab recomputes the p element
    // layers its first rows share with the previous slab, so every entry is
    // still summed over all its elements in element order); several patches
    // are assembled at once and must fit the budget.
    const HostPatch &h0 = P->hp[0];
    const int sd = h0.d - 1;   // slowest direction
    int64_t layer_elems = 1;
    for (int k = 0; k < sd; ++k) layer_elems *= h0.nel[k];
    const int nb0 = h0.d == 3 ? (h0.p[0] + 1) * (h0.p[0] + 1) * (h0.p[0] + 1) : (h0.p[0] + 1) * (h0.p[0] + 1);
    int slab_rows = 0;   // row layers per slab (single patch)
    if (use_sf) {
        slab_rows = 0;
    } else if (single) {
        const int64_t layer_bytes = layer_elems * ((int64_t)nb0 * nb0 + nb0) * (int64_t)sizeof(double);
        const int64_t fit = c->elem_budget / std::max<int64_t>(1, layer_bytes);
        const int free_layers = P->box_hi[sd] - P->box_lo[sd] + 1;
        slab_rows = (int)std::min<int64_t>(free_layers, fit - h0.p[0]);
        if (slab_rows < 1)
            return fail(c, GSX_ERR_BUDGET, "one slab of element matrices exceeds the element budget");
    } else if ((P->kdoubles + P->fdoubles) * (int64_t)sizeof(double) > c->elem_budget) {
        return fail(c, GSX_ERR_BUDGET, "element matrices of a multipatch assembly exceed the element budget "
                                       "(gsx_ctx_set_elem_budget)");
    }
    const int64_t kcap = use_sf ? 0
                         : single ? std::min<int64_t>(h0.nel[sd], slab_rows + h0.p[0]) * layer_elems * nb0 * nb0
                                  : P->kdoubles;
    const int64_t fcap = use_sf ? P->nelem * nb0
                         : single ? std::min<int64_t>(h0.nel[sd], slab_rows + h0.p[0]) * layer_elems * nb0
                                  : P->fdoubles;
    DevPtr Kbuf = dalloc(c, sizeof(double) * std::max<int64_t>(1, kcap));
    DevPtr Fbuf = dalloc(c, sizeof(double) * std::max<int64_t>(1, fcap));
    if (!Kbuf || !Fbuf) return fail(c, GSX_ERR_OOM, "device allocation (element matrices)");
    if (!single && P->geo) {
        KScope k(c, "geo_elem");
        for (int ip = 0; ip < P->npatch; ++ip)
            HIP_OK(c, gsx::launch_geo_elem(P->d, dp<PatchDev>(P->d_patches), dp<PatchDev>(P->geo->d_patches), ip, 0,
                                           P->hp[ip].nelem, c->gt, dp<double>(Kbuf), dp<double>(Fbuf),
                                           dp<double>(P->d_detmin), 0, 0, 0, s));
    } else if (!single) {
        for (auto &g : P->groups) {
            KScope k(c, g.d == 3 ? "elem3d" : "elem2d");
            HIP_OK(c, gsx::launch_elem(g.d, g.p, dp<PatchDev>(P->d_patches), dp<int>(g.d_gpatch),
                                       dp<int64_t>(g.d_gprefix), (int)g.patches.size(), g.nelem, c->gt,
                                       dp<double>(Kbuf), dp<double>(Fbuf), dp<double>(P->d_detmin), s));
        }
    }

    auto A = std::make_unique<gsx_matrix>();
    A->ctx = c;
    A->n = P->nfree;
    A->nnz = P->nnz;
    A->rowptr = P->d_rowptr;
    A->col = P->d_col;
    A->val = dalloc(c, sizeof(double) * std::max<int64_t>(1, P->nnz));
    DevPtr rhs_tmp;
    double *rhs = rhs_dev;
    if (!rhs) {
        rhs_tmp = dalloc(c, sizeof(double) * std::max(1, P->nfree));
        rhs = dp<double>(rhs_tmp);
    }
    DevPtr keep;
    if (mode == GSX_PATTERN_EXPR && !use_sf) keep = dalloc(c, std::max<int64_t>(1, P->nnz));
    if (!A->val || !rhs || (mode == GSX_PATTERN_EXPR && !use_sf && !keep))
        return fail(c, GSX_ERR_OOM, "device allocation (values)");
    if (single) {
        // structured operator filled by the same gather (unless CSR only)
        StencilOp &S = A->st;
        const HostPatch &h = P->hp[0];
        S.available = !csr_only;
        S.d = h.d;
        S.p = h.p[0];
        S.nx = P->grid_dims[0];
        S.ny = P->grid_dims[1];
        S.nz = P->grid_dims[2];
        S.ngrid = (int64_t)S.nx * S.ny * S.nz;
        for (int k = 0; k < h.d; ++k) S.m[k] = P->box_hi[k] - P->box_lo[k] + 1;
        const int W = 2 * S.p + 1;
        S.slots = h.d == 3 ? W * W * W : W * W;
        S.f2g = P->d_f2g;
        if (!csr_only) {
            S.s64 = dalloc(c, sizeof(double) * gsx::stencil_size(A->n, (int)S.slots));
            if (precision == GSX_MIXED)
                S.s32 = dalloc(c, sizeof(float) * gsx::stencil_size(A->n, gsx::stencil_half((int)S.slots)));
            if (!S.s64 || (precision == GSX_MIXED && !S.s32))
                return fail(c, GSX_ERR_OOM, "device allocation (structured operator)");
        }
        S.s64_ready = !csr_only;
        S.s32_ready = !csr_only && precision == GSX_MIXED;
        if (use_sf) {
            gsx::SfBox sb;
            const int p = h.p[0];
            for (int k = 0; k < 3; ++k) {
                sb.nb[k] = h.nb[k];
                sb.nel[k] = h.nel[k];
                sb.flo[k] = P->box_lo[k];
                sb.fhi[k] = P->box_hi[k];
                sb.m[k] = P->box_hi[k] - P->box_lo[k] + 1;
                sb.g[k] = h.nel[k] * (p + 1);
            }
            sb.toff[0] = 0;
            sb.toff[1] = (int64_t)sb.g[0] * (2 * (p + 1) + 1);
            sb.toff[2] = sb.toff[1] + (int64_t)sb.g[1] * (2 * (p + 1) + 1);
            sb.nty = (h.nb[1] + 1) / 2;
            sb.ntz = (h.nb[2] + 1) / 2;
            DevPtr tab = dalloc(c, sizeof(double) * gsx::sf_table_doubles(p, h.nel));
            DevPtr dq = dalloc(c, sizeof(double) * std::max<int64_t>(1, gsx::sf_dq_doubles(p, h.nel)));
            DevPtr s64tmp;
            double *s64 = dp<double>(S.s64);
            if (!s64) {
                s64tmp = dalloc(c, sizeof(double) * gsx::stencil_size(A->n, (int)S.slots));
                s64 = dp<double>(s64tmp);
            }
            DevPtr errf = dalloc(c, sizeof(int));
            if (!tab || !dq || !s64 || !errf) return fail(c, GSX_ERR_OOM, "device allocation (row assembly)");
            HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int), s));
            {
                KScope k(c, "sf_assemble");
                HIP_OK(c, gsx::launch_sf_assemble(p, dp<PatchDev>(P->d_patches), c->gt, dp<double>(tab), sb,
                                                  dp<double>(Fbuf), dp<double>(P->d_detmin), dp<double>(dq), s64,
                                                  P->nelem, s));
            }
            {
                KScope k(c, "sf_finish");
                HIP_OK(c, gsx::launch_sf_finish(p, sb, P->nfree, dp<int>(P->d_rowptr), s64,
                                                S.s32 ? dp<float>(S.s32) : nullptr, dp<double>(A->val), rhs,
                                                dp<double>(Fbuf), dp<int>(P->d_dofs),
                                                P->d_fixed ? dp<double>(P->d_fixed) : nullptr, dp<int>(errf), s));
            }
            int errh = 0;
            HIP_OK(c, hipMemcpyAsync(&errh, errf.get(), sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_OK(c, hipStreamSynchronize(s));
            if (errh) return fail(c, GSX_ERR_ARG, "row assembly: CSR row length does not match the stencil window");
        }
        // free rows are lexicographic over the free box: a range of row
        // layers along the slowest direction is a range of rows
        int64_t plane = 1;
        for (int k = 0; k < sd; ++k) plane *= P->box_hi[k] - P->box_lo[k] + 1;
        const Group &g = P->groups[0];
        for (int L0 = P->box_lo[sd]; !use_sf && L0 <= P->box_hi[sd]; L0 += slab_rows) {
            const int L1 = std::min(L0 + slab_rows, P->box_hi[sd] + 1);
            const int64_t E0 = std::max(0, L0 - h.p[0]), E1 = std::min(h.nel[sd], L1);
            if (P->geo) {
                KScope k(c, "geo_elem");
                // element matrices of the slab from offset 0 (the kernel
                // adds the patch offsets, as the range element kernels do)
                HIP_OK(c, gsx::launch_geo_elem(h.d, dp<PatchDev>(P->d_patches), dp<PatchDev>(P->geo->d_patches), 0,
                                               E0 * layer_elems, E1 * layer_elems, c->gt, dp<double>(Kbuf),
                                               dp<double>(Fbuf), dp<double>(P->d_detmin),
                                               E0 * layer_elems * nb0 * nb0, E0 * layer_elems * nb0, 0, s));
            } else {
                KScope k(c, h.d == 3 ? "elem3d" : "elem2d");
                HIP_OK(c, gsx::launch_elem_range(h.d, h.p[0], dp<PatchDev>(P->d_patches), dp<int>(g.d_gpatch),
                                                 dp<int64_t>(g.d_gprefix), 1, E0 * layer_elems,
                                                 E1 * layer_elems, c->gt, dp<double>(Kbuf), dp<double>(Fbuf),
                                                 dp<double>(P->d_detmin), s));
            }
            KScope k(c, "gather");
            const int r0 = (int)((L0 - P->box_lo[sd]) * plane), r1 = (int)((L1 - P->box_lo[sd]) * plane);
            HIP_OK(c, gsx::launch_gather_patch(h.d, h.p[0], P->box_lo, P->box_hi, h.nb, h.nel, P->grid_dims,
                                               P->nfree, r0, r1, E0 * layer_elems, dp<int>(P->d_f2g),
                                               dp<int>(P->d_rowptr), dp<double>(Kbuf), dp<double>(Fbuf),
                                               dp<double>(A->val), rhs, keep ? dp<unsigned char>(keep) : nullptr,
                                               S.s64 ? dp<double>(S.s64) : nullptr,
                                               S.s32 ? dp<float>(S.s32) : nullptr, dp<int>(P->d_dofs),
                                               P->d_fixed ? dp<double>(P->d_fixed) : nullptr, s));
        }
    } else {
        KScope k(c, "gather");
        HIP_OK(c, gsx::launch_gather(dp<PatchDev>(P->d_patches), dp<int>(P->d_occ_ptr),
                                     dp<int>(P->d_occ_patch), dp<int64_t>(P->d_occ_local), P->nfree,
                                     dp<int>(P->d_rowptr), dp<int>(P->d_col), dp<double>(Kbuf),
                                     dp<double>(Fbuf), dp<double>(A->val), rhs,
                                     keep ? dp<unsigned char>(keep) : nullptr, P->maxrow,
                                     P->d_fixed ? dp<double>(P->d_fixed) : nullptr, s));
    }
    // element matrices go back to the cache once the gathers have run, or
    // stay with the problem when asked for and formed in one pass
    if (keep_elements && kcap == P->kdoubles) {
        P->kept_K = std::move(Kbuf);
        P->kept_serial = serial;
    }
    Kbuf.reset();
    Fbuf.reset();
    double *dmin_slot = dp<double>(P->d_detmin) + std::max<int64_t>(1, P->nelem);
    HIP_OK(c, gsx::launch_min(dp<double>(P->d_detmin), P->nelem, dmin_slot, s));
    double mind = 0.0;
    HIP_OK(c, hipMemcpyAsync(&mind, dmin_slot, sizeof(double), hipMemcpyDeviceToHost, s));
    if (mode == GSX_PATTERN_EXPR && keep) {
        DevPtr dropped = dalloc(c, sizeof(unsigned long long));
        if (!dropped) return fail(c, GSX_ERR_OOM, "device allocation");
        HIP_OK(c, hipMemsetAsync(dropped.get(), 0, sizeof(unsigned long long), s));
        HIP_OK(c, gsx::launch_keep(nullptr, nullptr, nullptr, dp<unsigned char>(keep), P->nfree, P->nnz,
                                   nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                   dp<unsigned long long>(dropped), 0, s));
        unsigned long long ndrop = 0;
        HIP_OK(c, hipMemcpyAsync(&ndrop, dropped.get(), sizeof(ndrop), hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipStreamSynchronize(s));
        if (ndrop > 0) {
            // compaction: entries whose every element contribution was an
            // exact zero are not part of the gsExprAssembler pattern
            const int n = P->nfree;
            DevPtr counts = dalloc(c, sizeof(int) * (n + 1));
            DevPtr nrp = dalloc(c, sizeof(int) * (n + 1));
            const int64_t nn = P->nnz - (int64_t)ndrop;
            DevPtr ncol = dalloc(c, sizeof(int) * std::max<int64_t>(1, nn));
            DevPtr nval = dalloc(c, sizeof(double) * std::max<int64_t>(1, nn));
            if (!counts || !nrp || !ncol || !nval) return fail(c, GSX_ERR_OOM, "device allocation");
            HIP_OK(c, gsx::launch_keep(dp<int>(A->rowptr), dp<int>(A->col), dp<double>(A->val),
                                       dp<unsigned char>(keep), n, P->nnz, dp<int>(counts), dp<int>(nrp),
                                       dp<int>(ncol), dp<double>(nval), dp<long long>(P->d_scan_bsum),
                                       dp<long long>(P->d_scan_total), nullptr, 1, s));
            A->rowptr = nrp;
            A->col = ncol;
            A->val = nval;
            A->nnz = nn;
        }
    }
    HIP_OK(c, hipStreamSynchronize(s));
    P->min_det = mind;
    if (!(mind > 0.0)) {
        char buf[128];
        std::snprintf(buf, sizeof(buf),
                      "Jacobian determinant vanishes or changes sign in a patch (min oriented det J = %.6e)", mind);
        return fail(c, GSX_ERR_JACOBIAN, buf);
    }
    A->lpr = gsx::spmv_lpr(A->n ? (double)A->nnz / A->n : 1.0);
    A->mixed = precision == GSX_MIXED;
    A->src = P;
    A->src_serial = serial;
    if (precision == GSX_MIXED && !A->st.available) {
        A->val32 = dalloc(c, sizeof(float) * std::max<int64_t>(1, A->nnz));
        if (!A->val32) return fail(c, GSX_ERR_OOM, "device allocation (fp32 values)");
        KScope k(c, "convert_f32");
        HIP_OK(c, gsx::launch_f64_to_f32(A->nnz, dp<double>(A->val), dp<float>(A->val32), s));
    }
    *A_out = A.release();
    return GSX_OK;
}

gsx_status gsx_matrix_from_csr(gsx_ctx *c, int32_t n, const int32_t *rp, const int32_t *ci,
                               const double *v, int32_t with_fp32, gsx_matrix **out)
{
    if (!c || n < 0 || !rp || !out) return GSX_ERR_ARG;
    const int64_t nnz = rp[n];
    if (nnz < 0 || (nnz > 0 && (!ci || !v))) return fail(c, GSX_ERR_ARG, "invalid CSR arrays");
    for (int i = 0; i < n; ++i)
        if (rp[i + 1] < rp[i]) return fail(c, GSX_ERR_ARG, "row pointers must not decrease");
    for (int64_t k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= n) return fail(c, GSX_ERR_ARG, "column index out of range");
    hipStream_t s = c->stream;
    auto A = std::make_unique<gsx_matrix>();
    A->ctx = c;
    A->n = n;
    A->nnz = nnz;
    A->rowptr = dalloc(c, sizeof(int) * (n + 1));
    A->col = dalloc(c, sizeof(int) * std::max<int64_t>(1, nnz));
    A->val = dalloc(c, sizeof(double) * std::max<int64_t>(1, nnz));
    if (!A->rowptr || !A->col || !A->val) return fail(c, GSX_ERR_OOM, "device allocation (matrix)");
    HIP_OK(c, hipMemcpyAsync(A->rowptr.get(), rp, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (nnz) {
        HIP_OK(c, hipMemcpyAsync(A->col.get(), ci, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
        HIP_OK(c, hipMemcpyAsync(A->val.get(), v, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    }
    A->lpr = gsx::spmv_lpr(n ? (double)nnz / n : 1.0);
    A->mixed = with_fp32 != 0;
    if (A->mixed) {
        A->val32 = dalloc(c, sizeof(float) * std::max<int64_t>(1, nnz));
        if (!A->val32) return fail(c, GSX_ERR_OOM, "device allocation (fp32 values)");
        HIP_OK(c, gsx::launch_f64_to_f32(nnz, dp<double>(A->val), dp<float>(A->val32), s));
    }
    HIP_OK(c, hipStreamSynchronize(s));
    *out = A.release();
    return GSX_OK;
}

gsx_status gsx_ctx_release_cache(gsx_ctx *c)
{
    if (!c) return GSX_ERR_ARG;
    (void)hipStreamSynchronize(c->stream);
    c->mem->trim();
    return GSX_OK;
}

gsx_status gsx_matrix_destroy(gsx_matrix *A)
{
    if (!A) return GSX_ERR_ARG;
    (void)hipStreamSynchronize(A->ctx->stream);
    delete A;
    return GSX_OK;
}

gsx_status gsx_matrix_info(const gsx_matrix *A, int64_t *n, int64_t *nnz)
{
    if (!A) return GSX_ERR_ARG;
    if (n) *n = A->n;
    if (nnz) *nnz = A->nnz;
    return GSX_OK;
}

gsx_status gsx_matrix_download(const gsx_matrix *A, int32_t *rp, int32_t *ci, double *v)
{
    if (!A) return GSX_ERR_ARG;
    gsx_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    if (rp) HIP_OK(c, hipMemcpyAsync(rp, A->rowptr.get(), sizeof(int) * (A->n + 1), hipMemcpyDeviceToHost, s));
    if (ci) HIP_OK(c, hipMemcpyAsync(ci, A->col.get(), sizeof(int) * A->nnz, hipMemcpyDeviceToHost, s));
    if (v) HIP_OK(c, hipMemcpyAsync(v, A->val.get(), sizeof(double) * A->nnz, hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    return GSX_OK;
}

gsx_status gsx_matrix_device_ptrs(const gsx_matrix *A, const int32_t **rp, const int32_t **ci,
                                  const double **v)
{
    if (!A) return GSX_ERR_ARG;
    if (rp) *rp = dp<int>(A->rowptr);
    if (ci) *ci = dp<int>(A->col);
    if (v) *v = dp<double>(A->val);
    return GSX_OK;
}

gsx_status gsx_spmv(gsx_matrix *A, const double *x, double *y)
{
    if (!A || !x || !y) return GSX_ERR_ARG;
    gsx_ctx *c = A->ctx;
    KScope k(c, "spmv_f64");
    HIP_OK(c, gsx::launch_spmv_f64(A->lpr, A->n, dp<int>(A->rowptr), dp<int>(A->col),
                                   dp<double>(A->val), x, y, nullptr, nullptr, 0, nullptr, c->stream));
    return GSX_OK;
}

}  // extern "C"

static gsx_status cg_workspace(gsx_matrix *A, int maxit, bool mixed, bool grid = false)
{
    gsx_ctx *c = A->ctx;
    CGWork &w = A->ws;
    const int n = A->n;
    if (grid) {
        // grid-layout solver: its vectors live in the structured operator
        if (!w.state) w.state = dalloc(c, sizeof(CGState));
        if (w.hist_cap < maxit + 1) {
            w.hist = dalloc(c, sizeof(double) * (maxit + 1));
            w.hist_cap = maxit + 1;
        }
        if (!w.state || !w.hist) return fail(c, GSX_ERR_OOM, "device allocation (CG state)");
        return GSX_OK;
    }
    if (w.n != n || !w.r) {
        w = CGWork();
        w.n = n;
        w.nparts = std::max(gsx::spmv_blocks(n, A->lpr), (n + 255) / 256) + 1;
        w.r = dalloc(c, sizeof(double) * n);
        w.z = dalloc(c, sizeof(double) * n);
        w.p = dalloc(c, sizeof(double) * n);
        w.t = dalloc(c, sizeof(double) * n);
        w.parts = dalloc(c, sizeof(double) * 4 * w.nparts);
        w.state = dalloc(c, sizeof(CGState));
        if (!w.r || !w.z || !w.p || !w.t || !w.parts || !w.state)
            return fail(c, GSX_ERR_OOM, "device allocation (CG workspace)");
        if (!A->diag) {
            A->diag = dalloc(c, sizeof(double) * n);
            if (!A->diag) return fail(c, GSX_ERR_OOM, "device allocation (diag)");
            HIP_OK(c, gsx::launch_diag_f64(n, dp<int>(A->rowptr), dp<int>(A->col), dp<double>(A->val),
                                           dp<double>(A->diag), c->stream));
        }
    }
    if (w.hist_cap < maxit + 1) {
        w.hist = dalloc(c, sizeof(double) * (maxit + 1));
        if (!w.hist) return fail(c, GSX_ERR_OOM, "device allocation (history)");
        w.hist_cap = maxit + 1;
    }
    if (mixed && !w.r32) {
        w.r32 = dalloc(c, sizeof(float) * n);
        w.z32 = dalloc(c, sizeof(float) * n);
        w.p32 = dalloc(c, sizeof(float) * n);
        w.t32 = dalloc(c, sizeof(float) * n);
        w.xc32 = dalloc(c, sizeof(float) * n);
        w.diag32 = dalloc(c, sizeof(float) * n);
        if (!w.r32 || !w.z32 || !w.p32 || !w.t32 || !w.xc32 || !w.diag32)
            return fail(c, GSX_ERR_OOM, "device allocation (mixed workspace)");
        HIP_OK(c, gsx::launch_diag_f32(n, dp<int>(A->rowptr), dp<int>(A->col), dp<float>(A->val32),
                                       dp<float>(w.diag32), c->stream));
    }
    return GSX_OK;
}

static gsx_status read_state(gsx_ctx *c, const CGWork &w, CGState &h)
{
    HIP_OK(c, hipMemcpyAsync(&h, w.state.get(), sizeof(CGState), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return GSX_OK;
}

// Device buffers of the FP64 CSR CG: four vectors, the diagonal, partial
// sums (4 x nparts), the state and the residual history (maxit + 1).
struct CgBufs {
    double *r, *z, *p, *t;
    const double *diag;
    double *parts;
    int nparts;
    CGState *state;
    double *hist;
};

// FP64 preconditioned CG with the reference's step order and stopping rule.
// keep_hist: copy the residual history into A->hist.
static gsx_status cg_fp64_run(gsx_matrix *A, const CgBufs &B, int precond, double tau, const double *b, double *x,
                              double tol, int maxit, int *iters, double *relres, bool keep_hist)
{
    gsx_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    const int n = A->n;
    const int np = B.nparts;
    // partial counts: SpMV kernels write ns per-block partials, vector kernels nv
    const int ns = gsx::spmv_blocks(n, A->lpr), nv = (n + 255) / 256;
    double *P0 = B.parts, *P1 = P0 + np, *P2 = P1 + np;
    CGState *st = B.state;
    double *hist = B.hist;
    CGState h{};
    h.tol = tol;
    h.tau = tau;
    h.maxit = maxit;
    HIP_OK(c, hipMemcpyAsync(st, &h, sizeof(CGState), hipMemcpyHostToDevice, s));
    const int *rp = dp<int>(A->rowptr), *ci = dp<int>(A->col);
    const double *va = dp<double>(A->val);
    double *r = B.r, *z = B.z, *p = B.p, *t = B.t;
    const double *diag = B.diag;
    auto state = [&](CGState &out) -> gsx_status {
        HIP_OK(c, hipMemcpyAsync(&out, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipStreamSynchronize(s));
        return GSX_OK;
    };
    {
        KScope k(c, "cg_init");
        // r = b - A x ; ||r||^2 ; ||b||^2
        HIP_OK(c, gsx::launch_spmv_f64(A->lpr, n, rp, ci, va, x, r, b, P1, 2, nullptr, s));
        HIP_OK(c, gsx::launch_dot_f64(n, b, b, P0, nullptr, s));
        HIP_OK(c, gsx::launch_cg_scalar(0, P0, nv, P1, ns, st, hist, s));
    }
    GSX_TRY(state(h));
    if (h.done == 3) HIP_OK(c, gsx::launch_fill_f64(x, n, 0.0, s));
    if (!h.done) {
        if (precond == GSX_PREC_JACOBI) {
            HIP_OK(c, gsx::launch_jacobi_apply_f64(n, r, p, diag, tau, P0, st, s));
        } else {
            HIP_OK(c, hipMemcpyAsync(p, r, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            HIP_OK(c, gsx::launch_dot_f64(n, r, r, P0, st, s));
        }
        HIP_OK(c, gsx::launch_cg_scalar(1, P0, nv, nullptr, 0, st, hist, s));
        const int chunk = 16;
        for (int done_it = 0; done_it < maxit;) {
            const int k_end = std::min(maxit, done_it + chunk);
            for (int k = done_it; k < k_end; ++k) {
                {
                    KScope ks(c, "spmv_f64");
                    HIP_OK(c, gsx::launch_spmv_f64(A->lpr, n, rp, ci, va, p, t, nullptr, P0, 1, st, s));
                }
                KScope kv(c, "cg_vec");
                HIP_OK(c, gsx::launch_cg_scalar(2, P0, ns, nullptr, 0, st, hist, s));
                HIP_OK(c, gsx::launch_cg_update_f64(n, x, r, p, t, z,
                                                    precond == GSX_PREC_JACOBI ? diag : nullptr, P1, P2, st, s));
                if (precond == GSX_PREC_JACOBI) {
                    HIP_OK(c, gsx::launch_cg_scalar(3, P1, nv, P2, nv, st, hist, s));
                    HIP_OK(c, gsx::launch_cg_pupdate_f64(n, p, z, st, s));
                } else {
                    HIP_OK(c, gsx::launch_cg_scalar(3, P1, nv, P1, nv, st, hist, s));
                    HIP_OK(c, gsx::launch_cg_pupdate_f64(n, p, r, st, s));
                }
            }
            done_it = k_end;
            GSX_TRY(state(h));
            if (h.done) break;
        }
    }
    *iters = h.it;
    *relres = h.err;
    if (keep_hist) {
        A->hist.resize(h.it + 1);
        HIP_OK(c, hipMemcpyAsync(A->hist.data(), hist, sizeof(double) * (h.it + 1), hipMemcpyDeviceToHost, s));
    }
    HIP_OK(c, hipStreamSynchronize(s));
    if (h.done == 1 || h.done == 3) return GSX_OK;
    if (h.err < tol) return GSX_OK;
    return fail(c, GSX_ERR_NOT_CONVERGED, "CG reached the maximum number of iterations");
}

static gsx_status cg_fp64(gsx_matrix *A, int precond, double tau, const double *b, double *x,
                          double tol, int maxit, int *iters, double *relres)
{
    CGWork &w = A->ws;
    const CgBufs B{dp<double>(w.r), dp<double>(w.z), dp<double>(w.p), dp<double>(w.t), dp<double>(A->diag),
                   dp<double>(w.parts), w.nparts, dp<CGState>(w.state), dp<double>(w.hist)};
    return cg_fp64_run(A, B, precond, tau, b, x, tol, maxit, iters, relres, true);
}

// Mixed precision: CG iterations on the FP32 copy of A with FP64 reliable
// updates: every time the FP32 recurrence residual has dropped by a factor
// delta (or claims convergence), the FP32 correction is added to the FP64
// solution and the residual is recomputed as b - A x with the FP64 matrix.
// The stopping test uses that FP64 residual only, so the fixed point is the
// solution of the FP64 system.
static gsx_status cg_mixed(gsx_matrix *A, int precond, double tau, const double *b, double *x,
                           double tol, int maxit, int *iters, double *relres)
{
    gsx_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    CGWork &w = A->ws;
    const int n = A->n;
    const int np = w.nparts;
    // partial counts: SpMV kernels write ns per-block partials, vector kernels nv
    const int ns = gsx::spmv_blocks(n, A->lpr), nv = (n + 255) / 256;
    double *P0 = dp<double>(w.parts), *P1 = P0 + np, *P2 = P1 + np;
    CGState *st = dp<CGState>(w.state);
    CGState h{};
    const double delta = 0.01;
    h.tol = tol;
    h.tau = tau;
    h.maxit = maxit;
    h.rel_since_update = delta * delta;
    HIP_OK(c, hipMemcpyAsync(st, &h, sizeof(CGState), hipMemcpyHostToDevice, s));
    const int *rp = dp<int>(A->rowptr), *ci = dp<int>(A->col);
    const double *va = dp<double>(A->val);
    const float *va32 = dp<float>(A->val32);
    double *r = dp<double>(w.r);
    float *r32 = dp<float>(w.r32), *z32 = dp<float>(w.z32), *p32 = dp<float>(w.p32),
          *t32 = dp<float>(w.t32), *xc = dp<float>(w.xc32);
    const float *d32 = dp<float>(w.diag32);
    const bool jac = precond == GSX_PREC_JACOBI;
    {
        KScope k(c, "cg_init");
        HIP_OK(c, gsx::launch_spmv_f64(A->lpr, n, rp, ci, va, x, r, b, P1, 2, nullptr, s));
        HIP_OK(c, gsx::launch_dot_f64(n, b, b, P0, nullptr, s));
        HIP_OK(c, gsx::launch_cg_scalar(0, P0, nv, P1, ns, st, nullptr, s));
    }
    GSX_TRY(read_state(c, w, h));
    if (h.done == 3) HIP_OK(c, hipMemsetAsync(x, 0, sizeof(double) * n, s));

    // r32 = fp32(r64); z32 = M r32; rz = (r32, z32); the direction p32 is kept
    auto restart_from_r64 = [&]() -> gsx_status {
        HIP_OK(c, gsx::launch_f64_to_f32(n, r, r32, s));
        if (jac) {
            HIP_OK(c, gsx::launch_jacobi_apply_f32(n, r32, z32, d32, tau, P0, nullptr, s));
        } else {
            HIP_OK(c, hipMemcpyAsync(z32, r32, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
            HIP_OK(c, gsx::launch_dot_f32(n, r32, r32, P0, nullptr, s));
        }
        HIP_OK(c, gsx::launch_cg_scalar(1, P0, nv, nullptr, 0, st, nullptr, s));
        return GSX_OK;
    };
    if (!h.done) {
        HIP_OK(c, hipMemsetAsync(xc, 0, sizeof(float) * n, s));
        gsx_status e = restart_from_r64();
        if (e != GSX_OK) return e;
        HIP_OK(c, hipMemcpyAsync(p32, z32, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        int chunk = 8;
        while (true) {
            for (int k = 0; k < chunk; ++k) {
                {
                    KScope ks(c, "spmv_f32");
                    HIP_OK(c, gsx::launch_spmv_f32(A->lpr, n, rp, ci, va32, p32, t32, nullptr, P0, 1, st, s));
                }
                KScope kv(c, "cg_vec_f32");
                HIP_OK(c, gsx::launch_cg_scalar(2, P0, ns, nullptr, 0, st, nullptr, s));
                HIP_OK(c, gsx::launch_cg_update_f32(n, xc, r32, p32, t32, z32, jac ? d32 : nullptr, P1, P2, st, s));
                HIP_OK(c, gsx::launch_cg_scalar(4, P1, nv, jac ? P2 : P1, nv, st, nullptr, s));
                HIP_OK(c, gsx::launch_cg_pupdate_f32(n, p32, jac ? z32 : r32, st, s));
            }
            GSX_TRY(read_state(c, w, h));
            const bool due = h.need_update || h.done == 4 || h.done == 2 ||
                             h.rr < delta * delta * h.rr_ref || h.err < tol;
            if (due) {
                KScope kr(c, "reliable_update");
                const double rr_rec = h.rr;   // the FP32 recurrence's residual
                HIP_OK(c, gsx::launch_accum_correction(n, x, xc, s));
                HIP_OK(c, gsx::launch_spmv_f64(A->lpr, n, rp, ci, va, x, r, b, P1, 2, nullptr, s));
                HIP_OK(c, gsx::launch_cg_reliable(P1, ns, st, nullptr, s));
                GSX_TRY(read_state(c, w, h));
                if (h.done == 1 || h.done == 2) break;
                gsx_status e2 = restart_from_r64();
                if (e2 != GSX_OK) return e2;
                // a recurrence that has parted from the true residual (it
                // ran into its FP32 floor) leaves no usable direction
                if (rr_rec < gsx::kDirectionReset * h.rr)
                    HIP_OK(c, hipMemcpyAsync(p32, jac ? z32 : r32, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
            }
            if (h.it >= maxit) break;
        }
    }
    *iters = h.it;
    *relres = h.err;
    HIP_OK(c, hipStreamSynchronize(s));
    if (h.done == 1 || h.done == 3 || h.err < tol) return GSX_OK;
    return fail(c, GSX_ERR_NOT_CONVERGED, "mixed CG reached the maximum number of iterations");
}

static gsx_status cg_mg_entry(gsx_matrix *A, gsx_mg *mg, const double *b, double *x, double tol, int maxit,
                              int *iters, double *relres);

// ---------------------------------------------------------------------------
// structured operator + grid-layout CG
// ---------------------------------------------------------------------------
constexpr int kPartStride = 2048;

static gsx_status stencil_prepare(gsx_matrix *A, bool mixed)
{
    gsx_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    StencilOp &S = A->st;
    const int n = A->n;
    const int64_t K = S.slots, N = S.ngrid;
    KScope k(c, "stencil_build");
    if (!S.s64_ready) {
        const int64_t sz = gsx::stencil_size(n, (int)K);
        S.s64 = dalloc(c, sizeof(double) * sz);
        if (!S.s64) return fail(c, GSX_ERR_OOM, "device allocation (structured operator)");
        HIP_OK(c, gsx::launch_fill_f64(dp<double>(S.s64), sz, 0.0, s));
        HIP_OK(c, gsx::launch_csr_to_stencil(S.d, S.p, n, dp<int>(A->rowptr), dp<int>(A->col),
                                             dp<double>(A->val), dp<int>(S.f2g), S.nx, S.ny,
                                             dp<double>(S.s64), nullptr, s));
        S.s64_ready = true;
    }
    if (mixed && !S.s32_ready) {
        const int64_t sz = gsx::stencil_size(n, gsx::stencil_half((int)K));
        S.s32 = dalloc(c, sizeof(float) * sz);
        if (!S.s32) return fail(c, GSX_ERR_OOM, "device allocation (structured operator, fp32)");
        HIP_OK(c, gsx::launch_fill_f32(dp<float>(S.s32), sz, 0.0f, s));
        HIP_OK(c, gsx::launch_csr_to_stencil(S.d, S.p, n, dp<int>(A->rowptr), dp<int>(A->col),
                                             dp<double>(A->val), dp<int>(S.f2g), S.nx, S.ny, nullptr,
                                             dp<float>(S.s32), s));
        S.s32_ready = true;
    }
    if (!S.built) {
        S.diag64 = dalloc(c, sizeof(double) * N);
        S.xg = dalloc(c, sizeof(double) * N);
        S.rg = dalloc(c, sizeof(double) * N);
        S.bg = dalloc(c, sizeof(double) * N);
        S.zg = dalloc(c, sizeof(double) * N);
        S.pg = dalloc(c, sizeof(double) * N);
        S.tg = dalloc(c, sizeof(double) * N);
        S.parts = dalloc(c, sizeof(double) * 4 * kPartStride);
        if (!S.diag64 || !S.xg || !S.rg || !S.bg || !S.zg || !S.pg || !S.tg || !S.parts)
            return fail(c, GSX_ERR_OOM, "device allocation (grid workspace)");
        // diagonal = centre slot of the structured operator
        HIP_OK(c, gsx::launch_fill_f64(dp<double>(S.diag64), N, 1.0, s));
        HIP_OK(c, gsx::launch_stencil_diag(n, (int)K, dp<int>(S.f2g), dp<double>(S.s64), dp<double>(S.diag64), s));
        for (DevPtr *v : {&S.xg, &S.rg, &S.bg, &S.zg, &S.pg, &S.tg})
            HIP_OK(c, gsx::launch_fill_f64(dp<double>(*v), N, 0.0, s));
        S.built = true;
    }
    if (mixed && !S.r32) {
        S.diag32 = dalloc(c, sizeof(float) * N);
        S.r32 = dalloc(c, sizeof(float) * N);
        S.z32 = dalloc(c, sizeof(float) * N);
        S.p32 = dalloc(c, sizeof(float) * N);
        S.t32 = dalloc(c, sizeof(float) * N);
        S.xc32 = dalloc(c, sizeof(float) * N);
        if (!S.diag32 || !S.r32 || !S.z32 || !S.p32 || !S.t32 || !S.xc32)
            return fail(c, GSX_ERR_OOM, "device allocation (grid workspace, fp32)");
        HIP_OK(c, gsx::launch_grid_f64_to_f32(N, dp<double>(S.diag64), dp<float>(S.diag32), s));
        for (DevPtr *v : {&S.r32, &S.z32, &S.p32, &S.t32, &S.xc32})
            HIP_OK(c, gsx::launch_fill_f32(dp<float>(*v), N, 0.0f, s));
    }
    return GSX_OK;
}

// ===========================================================================
// Fast diagonalization (gsPatchPreconditionersCreator::fastDiagonalizationOp,
// src/gsSolver/gsPatchPreconditionersCreator.hpp:258-348) on the free box of
// a single patch: univariate mass M_k and stiffness beta K_k of each knot
// vector on the parameter line (Dirichlet ends removed, as
// eliminateDirichlet1D does), K_k q = lambda M_k q with Q_k^T M_k Q_k = I,
// and z = (Q_{d-1} x .. x Q_0) diag(1/(alpha + sum lambda)) (..)^T r, which
// is the exact inverse of alpha M + sum_k M x .. x beta K_k x .. x M.
// ===========================================================================
struct gsx_fd {
    gsx_ctx *ctx = nullptr;
    int kind = 0;                          // 0 fast diagonalization, 1 subspace corrected mass smoother
    int d = 0;
    int m[3] = {1, 1, 1};
    int64_t size = 0;
    double alpha = 0.0;
    DevPtr Q64[3], QT64[3], Q32[3], QT32[3], lam[3];
    DevPtr t64[2], t32[2];
    int grid[3] = {1, 1, 1};
    int64_t goff = 0, gs1 = 0, gs2 = 0;   // free box on the padded grid
    // subspace corrected mass smoother, per direction: S = B~ M~^-1 B~^T
    // (n x n), W = B_compl Q_c (n x cc) and its transpose, lambda_c
    int ct[3] = {0, 0, 0}, cc[3] = {0, 0, 0};
    DevPtr S64[3], S32[3], W64[3], W32[3], WT64[3], WT32[3], lc[3];
    DevPtr xin64;                          // copy of the input of an in-place apply (every type reads it)
    double beta = 1.0, sigma = 1.0, h = 1.0;
};

static hipError_t fd_mode(int m0, int m1, int m2, int dir, const double *M, const double *x, int64_t xo,
                          int64_t xs1, int64_t xs2, double *y, int64_t yo, int64_t ys1, int64_t ys2,
                          const double *l0, const double *l1, const double *l2, double alpha,
                          const gsx::CGState *st, hipStream_t s)
{
    return gsx::launch_fd_mode_f64(m0, m1, m2, dir, M, x, xo, xs1, xs2, y, yo, ys1, ys2, l0, l1, l2, alpha, st, s);
}

static hipError_t fd_mode(int m0, int m1, int m2, int dir, const float *M, const float *x, int64_t xo,
                          int64_t xs1, int64_t xs2, float *y, int64_t yo, int64_t ys1, int64_t ys2,
                          const double *l0, const double *l1, const double *l2, double alpha,
                          const gsx::CGState *st, hipStream_t s)
{
    return gsx::launch_fd_mode_f32(m0, m1, m2, dir, M, x, xo, xs1, xs2, y, yo, ys1, ys2, l0, l1, l2, alpha, st, s);
}

static hipError_t tensor_mode(int o0, int o1, int o2, int dir, int m_in, const double *M, const double *x,
                              int64_t xo, int64_t xs1, int64_t xs2, double *y, int64_t yo, int64_t ys1, int64_t ys2,
                              int scale, double alpha, const double *l0, const double *l1, const double *l2,
                              int accumulate, const gsx::CGState *st, hipStream_t s)
{
    return gsx::launch_tensor_mode_f64(o0, o1, o2, dir, m_in, M, x, xo, xs1, xs2, y, yo, ys1, ys2, scale, alpha, l0,
                                       l1, l2, accumulate, st, s);
}

static hipError_t tensor_mode(int o0, int o1, int o2, int dir, int m_in, const float *M, const float *x,
                              int64_t xo, int64_t xs1, int64_t xs2, float *y, int64_t yo, int64_t ys1, int64_t ys2,
                              int scale, double alpha, const double *l0, const double *l1, const double *l2,
                              int accumulate, const gsx::CGState *st, hipStream_t s)
{
    return gsx::launch_tensor_mode_f32(o0, o1, o2, dir, m_in, M, x, xo, xs1, xs2, y, yo, ys1, ys2, scale, alpha, l0,
                                       l1, l2, accumulate, st, s);
}

// Subspace corrected mass smoother (subspaceCorrectedMassSmootherOp,
// src/gsSolver/gsPatchPreconditionersCreator.hpp:574-758): the sum over the
// subspace types J (directions with the boundary complement; the others
// with the interior space) of T_J C_J^-1 T_J^T.  Per type: for k in J
// reduce with W_k^T, divide by s + sum_{k in J} lambda_k (s = alpha + beta
// (d - |J|) / (sigma h^2)), expand with W_k; for k not in J apply S_k (the
// interior type J = {} divides by s once).  Types missing a factor are
// skipped; the first type writes y, the others add to it.
template <typename T>
static hipError_t scms_apply(const gsx_fd *F, const T *x, int64_t xo, int64_t xs1, int64_t xs2, T *y, int64_t yo,
                             int64_t ys1, int64_t ys2, const gsx::CGState *st, hipStream_t s)
{
    constexpr bool f64 = std::is_same<T, double>::value;
    T *buf[2] = {static_cast<T *>(f64 ? F->t64[0].get() : F->t32[0].get()),
                 static_cast<T *>(f64 ? F->t64[1].get() : F->t32[1].get())};
    const int D = F->d;
    auto mat = [&](const DevPtr (&a64)[3], const DevPtr (&a32)[3], int k) {
        return static_cast<const T *>(f64 ? a64[k].get() : a32[k].get());
    };
    bool first = true;
    for (int J = 0; J < (1 << D); ++J) {
        bool skip = false;
        int nj = 0;
        for (int k = 0; k < D; ++k) {
            const bool inJ = (J >> k) & 1;
            nj += inJ;
            if (inJ ? F->cc[k] == 0 : F->ct[k] == 0) skip = true;
        }
        if (skip) continue;
        const double sd = F->alpha + F->beta * (double)(D - nj) / (F->sigma * F->h * F->h);
        struct Step { int dir; const T *M; int m_in, m_out; bool scale; };
        Step steps[6];
        int ns = 0, last_reduce = -1;
        for (int k = 0; k < D; ++k)
            if ((J >> k) & 1) {
                last_reduce = ns;
                steps[ns++] = {k, mat(F->W64, F->W32, k), F->m[k], F->cc[k], false};
            }
        for (int k = 0; k < D; ++k)
            if ((J >> k) & 1) steps[ns++] = {k, mat(F->WT64, F->WT32, k), F->cc[k], F->m[k], false};
        for (int k = 0; k < D; ++k)
            if (!((J >> k) & 1)) steps[ns++] = {k, mat(F->S64, F->S32, k), F->m[k], F->m[k], false};
        steps[last_reduce >= 0 ? last_reduce : 0].scale = true;
        const double *lc[3] = {nullptr, nullptr, nullptr};
        for (int k = 0; k < D; ++k)
            if ((J >> k) & 1) lc[k] = dp<double>(F->lc[k]);
        int cur[3] = {F->m[0], F->m[1], D == 3 ? F->m[2] : 1};
        const T *in = x;
        int64_t io = xo, is1 = xs1, is2 = xs2;
        for (int t = 0; t < ns; ++t) {
            const Step &S = steps[t];
            int out[3] = {cur[0], cur[1], cur[2]};
            out[S.dir] = S.m_out;
            const bool last = t == ns - 1;
            T *o = last ? y : buf[t & 1];
            const int64_t oo = last ? yo : 0, os1 = last ? ys1 : out[0], os2 = last ? ys2 : (int64_t)out[0] * out[1];
            const hipError_t e = tensor_mode(out[0], out[1], out[2], S.dir, S.m_in, S.M, in, io, is1, is2, o, oo, os1,
                                             os2, S.scale ? 1 : 0, sd, lc[0], lc[1], lc[2], last && !first ? 1 : 0,
                                             st, s);
            if (e != hipSuccess) return e;
            in = o;
            io = oo;
            is1 = os1;
            is2 = os2;
            cur[0] = out[0];
            cur[1] = out[1];
            cur[2] = out[2];
        }
        first = false;
    }
    return first ? hipErrorInvalidValue : hipSuccess;
}

// y = FD(x): element (a,b,c) of x at xo + a + b xs1 + c xs2 (likewise y); the
// 2d mode products alternate between the two compact scratch tensors.
template <typename T>
static hipError_t fd_apply(const gsx_fd *F, const T *x, int64_t xo, int64_t xs1, int64_t xs2, T *y, int64_t yo,
                           int64_t ys1, int64_t ys2, const gsx::CGState *st, hipStream_t s)
{
    if (F->kind == 1) return scms_apply<T>(F, x, xo, xs1, xs2, y, yo, ys1, ys2, st, s);
    constexpr bool f64 = std::is_same<T, double>::value;
    T *buf[2] = {static_cast<T *>(f64 ? F->t64[0].get() : F->t32[0].get()),
                 static_cast<T *>(f64 ? F->t64[1].get() : F->t32[1].get())};
    const int D = F->d, m0 = F->m[0], m1 = F->m[1], m2 = F->m[2];
    const double *l0 = dp<double>(F->lam[0]), *l1 = dp<double>(F->lam[1]);
    const double *l2 = D == 3 ? dp<double>(F->lam[2]) : nullptr;
    const T *in = x;
    int64_t io = xo, is1 = xs1, is2 = xs2;
    for (int k = 0; k < 2 * D; ++k) {
        const bool forward = k < D;   // Q^T along 0..D-1, then Q along D-1..0
        const int dir = forward ? k : 2 * D - 1 - k;
        const T *M = static_cast<const T *>(forward ? (f64 ? F->Q64[dir].get() : F->Q32[dir].get())
                                                    : (f64 ? F->QT64[dir].get() : F->QT32[dir].get()));
        const bool last = k == 2 * D - 1;
        T *out = last ? y : buf[k & 1];
        const int64_t oo = last ? yo : 0, os1 = last ? ys1 : m0, os2 = last ? ys2 : (int64_t)m0 * m1;
        const hipError_t e = fd_mode(m0, m1, m2, dir, M, in, io, is1, is2, out, oo, os1, os2,
                                     k == D - 1 ? l0 : nullptr, l1, l2, F->alpha, st, s);
        if (e != hipSuccess) return e;
        in = out;
        io = oo;
        is1 = os1;
        is2 = os2;
    }
    return hipSuccess;
}

// Grid-layout CG.  mixed == false: FP64 with the reference's recurrences and
// stopping rule.  mixed == true: FP32 recurrences on the FP32 copy of the
// operator with FP64 reliable updates (see cg_mixed).
static gsx_status cg_grid(gsx_matrix *A, int precond, double tau, const double *b, double *x,
                          double tol, int maxit, bool mixed, int *iters, double *relres, const gsx_fd *fd = nullptr)
{
    gsx_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    StencilOp &S = A->st;
    CGWork &w = A->ws;
    const int n = A->n;
    const int64_t N = S.ngrid;
    const int nsp = gsx::sspmv_blocks(n), nv = gsx::vec_blocks(N);
    double *P0 = dp<double>(S.parts), *P1 = P0 + kPartStride, *P2 = P1 + kPartStride;
    CGState *st = dp<CGState>(w.state);
    double *hist = mixed ? nullptr : dp<double>(w.hist);
    const bool jac = precond == GSX_PREC_JACOBI;
    const int *f2g = dp<int>(S.f2g);
    const int sy = S.nx, sz = S.nx * S.ny;
    double *xg = dp<double>(S.xg), *rg = dp<double>(S.rg), *bg = dp<double>(S.bg);
    double *zg = dp<double>(S.zg), *pg = dp<double>(S.pg), *tg = dp<double>(S.tg);
    const double *sv = dp<double>(S.s64), *dg = dp<double>(S.diag64);
    CGState h{};
    h.tol = tol;
    h.tau = tau;
    h.maxit = maxit;
    h.rel_since_update = 1e-4;   // mixed: reliable update every 1e-2 of residual reduction
    HIP_OK(c, hipMemcpyAsync(st, &h, sizeof(CGState), hipMemcpyHostToDevice, s));
    {
        KScope k(c, "cg_init");
        HIP_OK(c, gsx::launch_scatter_f64_to_grid(n, f2g, b, bg, s));
        HIP_OK(c, gsx::launch_scatter_f64_to_grid(n, f2g, x, xg, s));
        HIP_OK(c, gsx::launch_sspmv_f64(S.d, S.p, 2, n, f2g, sy, sz, sv, xg, rg, bg, P1, nullptr, 0, s));
        HIP_OK(c, gsx::launch_cg2_precond_f64((int)N, bg, zg, nullptr, 1.0, P0, s));   // ||b||^2
        HIP_OK(c, gsx::launch_cg_scalar(0, P0, nv, P1, nsp, st, hist, s));
    }
    GSX_TRY(read_state(c, w, h));
    if (h.done == 3) HIP_OK(c, gsx::launch_fill_f64(x, n, 0.0, s));
    int k = 0;
    // the fast diagonalization applies between the update (which then only
    // moves x and r) and the direction, followed by the (r, z) partials
    auto fd64 = [&](const gsx::CGState *sp, double *part) -> gsx_status {
        KScope kf(c, "fd_apply");
        HIP_OK(c, fd_apply<double>(fd, rg, fd->goff, fd->gs1, fd->gs2, zg, fd->goff, fd->gs1, fd->gs2, sp, s));
        HIP_OK(c, gsx::launch_cg2_dot_f64((int)N, rg, zg, part, sp, s));
        return GSX_OK;
    };
    if (!h.done && !mixed) {
        if (fd) {
            GSX_TRY(fd64(nullptr, P0));
        } else {
            HIP_OK(c, gsx::launch_cg2_precond_f64((int)N, rg, zg, jac ? dg : nullptr, tau, P0, s));
        }
        HIP_OK(c, gsx::launch_set_rz(P0, nv, st, 0, s));
        HIP_OK(c, hipMemcpyAsync(pg, zg, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
        const int chunk = 16;
        while (k < maxit) {
            const int kend = std::min(maxit, k + chunk);
            for (; k < kend; ++k) {
                {
                    KScope ks(c, "spmv_f64");
                    HIP_OK(c, gsx::launch_sspmv_f64(S.d, S.p, 1, n, f2g, sy, sz, sv, pg, tg, nullptr, P0, st, k, s));
                }
                {
                    KScope kv(c, "cg_vec");
                    HIP_OK(c, gsx::launch_cg2_update_f64((int)N, xg, rg, pg, tg, zg, jac && !fd ? dg : nullptr, tau,
                                                         P0, nsp, P1, P2, st, k, s));
                }
                if (fd) GSX_TRY(fd64(st, P2));
                KScope kv(c, "cg_vec");
                HIP_OK(c, gsx::launch_cg2_direction_f64((int)N, pg, zg, P1, P2, nv, st, k, hist, s));
            }
            GSX_TRY(read_state(c, w, h));
            if (h.done) break;
        }
        HIP_OK(c, gsx::launch_gather_f64_from_grid(n, f2g, xg, x, s));
    } else if (!h.done && mixed) {
        float *r32 = dp<float>(S.r32), *z32 = dp<float>(S.z32), *p32 = dp<float>(S.p32);
        float *t32 = dp<float>(S.t32), *xc = dp<float>(S.xc32);
        const float *sv32 = dp<float>(S.s32), *d32 = dp<float>(S.diag32);
        auto fd32 = [&](const gsx::CGState *sp, double *part) -> gsx_status {
            KScope kf(c, "fd_apply_f32");
            HIP_OK(c, fd_apply<float>(fd, r32, fd->goff, fd->gs1, fd->gs2, z32, fd->goff, fd->gs1, fd->gs2, sp, s));
            HIP_OK(c, gsx::launch_cg2_dot_f32((int)N, r32, z32, part, sp, s));
            return GSX_OK;
        };
        auto restart = [&](int slot) -> gsx_status {
            HIP_OK(c, gsx::launch_grid_f64_to_f32(N, rg, r32, s));
            if (fd) {
                GSX_TRY(fd32(nullptr, P0));
            } else {
                HIP_OK(c, gsx::launch_cg2_precond_f32((int)N, r32, z32, jac ? d32 : nullptr, tau, P0, s));
            }
            HIP_OK(c, gsx::launch_set_rz(P0, nv, st, slot, s));
            return GSX_OK;
        };
        HIP_OK(c, gsx::launch_fill_f32(xc, N, 0.0f, s));
        GSX_TRY(restart(0));
        HIP_OK(c, hipMemcpyAsync(p32, z32, sizeof(float) * N, hipMemcpyDeviceToDevice, s));
        const int chunk = 8;
        while (true) {
            for (int j = 0; j < chunk; ++j, ++k) {
                {
                    KScope ks(c, "spmv_f32");
                    HIP_OK(c, gsx::launch_sspmv_f32(S.d, S.p, 1, n, f2g, sy, sz, S.m[0], S.m[1], S.m[2], sv32, p32,
                                                    t32, nullptr, P0, st, k, s));
                }
                {
                    KScope kv(c, "cg_vec_f32");
                    HIP_OK(c, gsx::launch_cg2_update_f32((int)N, xc, r32, p32, t32, z32, jac && !fd ? d32 : nullptr,
                                                         tau, P0, nsp, P1, P2, st, k, s));
                }
                if (fd) GSX_TRY(fd32(st, P2));
                KScope kv(c, "cg_vec_f32");
                HIP_OK(c, gsx::launch_cg2_direction_f32_mixed((int)N, p32, z32, P1, P2, nv, st, k, s));
            }
            GSX_TRY(read_state(c, w, h));
            k = h.it;
            if (h.need_update || h.done == 2 || h.err < tol) {
                KScope kr(c, "reliable_update");
                const double rr_rec = h.rr;   // the FP32 recurrence's residual
                HIP_OK(c, gsx::launch_accum_grid(N, xg, xc, s));
                HIP_OK(c, gsx::launch_sspmv_f64(S.d, S.p, 2, n, f2g, sy, sz, sv, xg, rg, bg, P1, nullptr, 0, s));
                HIP_OK(c, gsx::launch_cg_reliable(P1, nsp, st, nullptr, s));
                GSX_TRY(read_state(c, w, h));
                if (h.done == 1 || h.done == 2) break;
                GSX_TRY(restart(k));
                // a recurrence that has parted from the true residual (it
                // ran into its FP32 floor) leaves no usable direction
                if (rr_rec < gsx::kDirectionReset * h.rr)
                    HIP_OK(c, hipMemcpyAsync(p32, z32, sizeof(float) * N, hipMemcpyDeviceToDevice, s));
            }
            if (h.it >= maxit) break;
        }
        HIP_OK(c, gsx::launch_gather_f64_from_grid(n, f2g, xg, x, s));
    }
    *iters = h.it;
    *relres = h.err;
    if (!mixed) {
        A->hist.resize(h.it + 1);
        HIP_OK(c, hipMemcpyAsync(A->hist.data(), hist, sizeof(double) * (h.it + 1), hipMemcpyDeviceToHost, s));
    }
    HIP_OK(c, hipStreamSynchronize(s));
    if (h.done == 1 || h.done == 3 || h.err < tol) return GSX_OK;
    return fail(c, GSX_ERR_NOT_CONVERGED, "CG reached the maximum number of iterations");
}

extern "C" {

gsx_status gsx_matrix_set_operator(gsx_matrix *A, int32_t structured)
{
    if (!A) return GSX_ERR_ARG;
    A->use_stencil = structured != 0;
    return GSX_OK;
}

gsx_status gsx_cg_solve(gsx_matrix *A, int32_t precond, double tau, gsx_mg *mg, int32_t precision,
                        const double *b, double *x, double tol, int32_t maxit, int32_t *iters,
                        double *relres)
{
    if (!A || !b || !x || maxit < 0) return GSX_ERR_ARG;
    gsx_ctx *c = A->ctx;
    if (A->n == 0) {   // no free DOFs: nothing to iterate (gsIterativeSolver::solve with b = 0)
        if (iters) *iters = 0;
        if (relres) *relres = 0.0;
        A->hist.assign(1, 0.0);
        return GSX_OK;
    }
    if (precond == GSX_PREC_MG || mg) {
        if (!mg) return fail(c, GSX_ERR_ARG, "multigrid preconditioner needs an gsx_mg");
        int it = 0;
        double rr = 0.0;
        gsx_status e = cg_mg_entry(A, mg, b, x, tol, maxit, &it, &rr);
        if (iters) *iters = it;
        if (relres) *relres = rr;
        return e;
    }
    if (precond != GSX_PREC_NONE && precond != GSX_PREC_JACOBI)
        return fail(c, GSX_ERR_ARG, "unknown preconditioner");
    if (precision == GSX_MIXED && !A->mixed)
        return fail(c, GSX_ERR_ARG, "mixed precision needs a matrix assembled with GSX_MIXED");
    const bool mixed = precision == GSX_MIXED;
    const bool grid = A->st.available && A->use_stencil;
    if (mixed && !grid && !A->val32) {
        // CSR mixed path on a single-patch matrix: FP32 copy made on demand
        A->val32 = dalloc(c, sizeof(float) * std::max<int64_t>(1, A->nnz));
        if (!A->val32) return fail(c, GSX_ERR_OOM, "device allocation (fp32 values)");
        KScope k(c, "convert_f32");
        HIP_OK(c, gsx::launch_f64_to_f32(A->nnz, dp<double>(A->val), dp<float>(A->val32), c->stream));
    }
    gsx_status e = cg_workspace(A, std::max(1, maxit), mixed && !grid, grid);
    if (e != GSX_OK) return e;
    int it = 0;
    double rr = 0.0;
    if (grid) {
        e = stencil_prepare(A, mixed);
        if (e != GSX_OK) return e;
        e = cg_grid(A, precond, tau, b, x, tol, maxit, mixed, &it, &rr);
    } else {
        e = mixed ? cg_mixed(A, precond, tau, b, x, tol, maxit, &it, &rr)
                  : cg_fp64(A, precond, tau, b, x, tol, maxit, &it, &rr);
    }
    if (iters) *iters = it;
    if (relres) *relres = rr;
    return e;
}

// gsx_cg_solve's FP64 CSR CG (GSX_PREC_NONE or GSX_PREC_JACOBI) with its
// workspace -- the four vectors, the diagonal, the partial sums, the state
// and the residual history -- in `scratch` instead of device allocations.
// Layout (256-byte aligned pieces): r, z, p, t, diag (n doubles each),
// partials (4 x nparts), history (maxit + 1), state.
gsx_status gsx_cg_solve_noalloc(gsx_matrix *A, int32_t precond, double tau, const double *b, double *x,
                                double tol, int32_t maxit, int32_t *iters, double *relres)
{
    if (!A || !b || !x || maxit < 0) return GSX_ERR_ARG;
    gsx_ctx *c = A->ctx;
    if (precond != GSX_PREC_NONE && precond != GSX_PREC_JACOBI)
        return fail(c, GSX_ERR_ARG, "allocation-free CG: GSX_PREC_NONE or GSX_PREC_JACOBI only");
    if (!c->scratch) return fail(c, GSX_ERR_HIP, "allocation-free CG: workspace address not set up");
    const int n = A->n;
    if (n == 0) {
        if (iters) *iters = 0;
        if (relres) *relres = 0.0;
        return GSX_OK;
    }
    const int nparts = std::max(gsx::spmv_blocks(n, A->lpr), (n + 255) / 256) + 1;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t vec = sizeof(double) * (size_t)n;
    const size_t o_r = take(vec), o_z = take(vec), o_p = take(vec), o_t = take(vec), o_d = take(vec);
    const size_t o_parts = take(sizeof(double) * 4 * (size_t)nparts);
    const size_t o_hist = take(sizeof(double) * ((size_t)maxit + 1));
    const size_t o_state = take(sizeof(CGState));
    if (off > kScratchBytes) {
        char buf[200];
        std::snprintf(buf, sizeof(buf),
                      "allocation-free CG: %d rows with maxit %d need %zu bytes, more than the %zu-byte workspace",
                      n, maxit, off, kScratchBytes);
        return fail(c, GSX_ERR_ARG, buf);
    }
    unsigned char *base = static_cast<unsigned char *>(c->scratch);
    double *diag = reinterpret_cast<double *>(base + o_d);
    if (precond == GSX_PREC_JACOBI)
        HIP_OK(c, gsx::launch_diag_f64(n, dp<int>(A->rowptr), dp<int>(A->col), dp<double>(A->val), diag, c->stream));
    const CgBufs B{reinterpret_cast<double *>(base + o_r), reinterpret_cast<double *>(base + o_z),
                   reinterpret_cast<double *>(base + o_p), reinterpret_cast<double *>(base + o_t), diag,
                   reinterpret_cast<double *>(base + o_parts), nparts, reinterpret_cast<CGState *>(base + o_state),
                   reinterpret_cast<double *>(base + o_hist)};
    int it = 0;
    double rr = 0.0;
    const gsx_status e = cg_fp64_run(A, B, precond, tau, b, x, tol, maxit, &it, &rr, false);
    if (iters) *iters = it;
    if (relres) *relres = rr;
    return e;
}

gsx_status gsx_fd_create(gsx_problem *P, double alpha, double beta, gsx_fd **out)
{
    if (!P || !out) return GSX_ERR_ARG;
    gsx_ctx *c = P->ctx;
    hipStream_t s = c->stream;
    if (P->npatch != 1 || !P->d_f2g)
        return fail(c, GSX_ERR_ARG, "fast diagonalization needs a single-patch problem whose free functions form a box");
    if (!(beta > 0.0) || !(alpha >= 0.0))
        return fail(c, GSX_ERR_ARG, "fast diagonalization needs beta > 0 and alpha >= 0");
    const HostPatch &h = P->hp[0];
    auto F = std::make_unique<gsx_fd>();
    F->ctx = c;
    F->d = h.d;
    F->alpha = alpha;
    double lmin = alpha;
    std::vector<std::vector<double>> hold64;   // host sources of the uploads, alive until the sync
    std::vector<std::vector<float>> hold32;
    auto up64 = [&](DevPtr &dst, std::vector<double> v) -> bool {
        dst = dalloc(c, sizeof(double) * std::max<size_t>(1, v.size()));
        if (!dst) return false;
        hold64.push_back(std::move(v));
        return hipMemcpyAsync(dst.get(), hold64.back().data(), sizeof(double) * hold64.back().size(),
                              hipMemcpyHostToDevice, s) == hipSuccess;
    };
    auto up32 = [&](DevPtr &dst, const std::vector<double> &v) -> bool {
        dst = dalloc(c, sizeof(float) * std::max<size_t>(1, v.size()));
        if (!dst) return false;
        hold32.emplace_back(v.begin(), v.end());
        return hipMemcpyAsync(dst.get(), hold32.back().data(), sizeof(float) * hold32.back().size(),
                              hipMemcpyHostToDevice, s) == hipSuccess;
    };
    for (int k = 0; k < h.d; ++k) {
        const int p = h.p[k], n = h.nb[k], lo = P->box_lo[k], m = P->box_hi[k] - lo + 1;
        if (m <= 0) return fail(c, GSX_ERR_ARG, "fast diagonalization: no free functions in a direction");
        // univariate matrices with the assembly's p+1 Gauss points per span
        std::vector<double> gx(p + 1), gw(p + 1), K, M;
        gauss_rule(p + 1, gx.data(), gw.data());
        gsx::dense::bspline_1d_matrices(P->knots_h[k], p, gx, gw, K, M);
        std::vector<double> Kf((size_t)m * m), Mf((size_t)m * m), lam, Q;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                Kf[(size_t)i * m + j] = beta * K[(size_t)(lo + i) * n + (lo + j)];
                Mf[(size_t)i * m + j] = M[(size_t)(lo + i) * n + (lo + j)];
            }
        if (!gsx::dense::gen_sym_eig(m, Kf, Mf, lam, Q))
            return fail(c, GSX_ERR_ARG, "fast diagonalization: the univariate eigenproblem failed");
        lmin += lam[0];
        std::vector<double> QT((size_t)m * m);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) QT[(size_t)j * m + i] = Q[(size_t)i * m + j];
        F->m[k] = m;
        if (!up32(F->Q32[k], Q) || !up32(F->QT32[k], QT) || !up64(F->Q64[k], std::move(Q)) ||
            !up64(F->QT64[k], std::move(QT)) || !up64(F->lam[k], std::move(lam)))
            return fail(c, GSX_ERR_OOM, "device allocation (fast diagonalization)");
    }
    if (!(lmin > 0.0))
        return fail(c, GSX_ERR_ARG, "fast diagonalization is singular (no Dirichlet end in any direction and alpha = 0)");
    F->size = (int64_t)F->m[0] * F->m[1] * F->m[2];
    if (F->size != P->nfree) return fail(c, GSX_ERR_ARG, "fast diagonalization: free box does not match the problem");
    for (int j = 0; j < 2; ++j) {
        F->t64[j] = dalloc(c, sizeof(double) * F->size);
        F->t32[j] = dalloc(c, sizeof(float) * F->size);
        if (!F->t64[j] || !F->t32[j]) return fail(c, GSX_ERR_OOM, "device allocation (fast diagonalization)");
    }
    const int pg = h.p[0], pz = h.d == 3 ? h.p[0] : 0;
    for (int k = 0; k < 3; ++k) F->grid[k] = P->grid_dims[k];
    F->gs1 = P->grid_dims[0];
    F->gs2 = (int64_t)P->grid_dims[0] * P->grid_dims[1];
    F->goff = (P->box_lo[0] + pg) + F->gs1 * (P->box_lo[1] + pg) + F->gs2 * (P->box_lo[2] + pz);
    HIP_OK(c, hipStreamSynchronize(s));
    *out = F.release();
    return GSX_OK;
}

// Free functions of a single patch as a box: degree, function count, knot
// vector and free index range per direction.
struct ScmsBox {
    int d = 0;
    int p[3] = {0, 0, 0}, nb[3] = {1, 1, 1}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    std::vector<double> knots[3];
};

static bool scms_box_of(const gsx_problem *P, ScmsBox &B)
{
    if (!P || P->npatch != 1 || !P->d_f2g) return false;
    const HostPatch &hp = P->hp[0];
    B.d = hp.d;
    for (int k = 0; k < hp.d; ++k) {
        B.p[k] = hp.p[k];
        B.nb[k] = hp.nb[k];
        B.lo[k] = P->box_lo[k];
        B.hi[k] = P->box_hi[k];
        B.knots[k] = P->knots_h[k];
    }
    return true;
}

// Factors of the subspace corrected mass smoother on the free box (compact
// vectors, direction 0 fastest); the grid layout fields are left unset.
static gsx_status scms_build(gsx_ctx *c, const ScmsBox &B, double sigma, double alpha, double beta,
                             std::unique_ptr<gsx_fd> &out)
{
    hipStream_t s = c->stream;
    if (!(beta != 0.0) || !(sigma > 0.0) || !(alpha >= 0.0))
        return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother needs beta != 0, sigma > 0, alpha >= 0");
    auto F = std::make_unique<gsx_fd>();
    F->ctx = c;
    F->kind = 1;
    F->d = B.d;
    F->alpha = alpha;
    F->beta = beta;
    F->sigma = sigma;
    // h: the smallest knot span of the basis (gsBasis::getMinCellLength)
    double h = 1e300;
    for (int k = 0; k < B.d; ++k) {
        const std::vector<double> &U = B.knots[k];
        for (size_t i = 0; i + 1 < U.size(); ++i)
            if (U[i + 1] > U[i]) h = std::min(h, U[i + 1] - U[i]);
    }
    F->h = h;
    std::vector<std::vector<double>> hold64;
    std::vector<std::vector<float>> hold32;
    auto up = [&](DevPtr &d64, DevPtr &d32, std::vector<double> v) -> bool {
        d64 = dalloc(c, sizeof(double) * std::max<size_t>(1, v.size()));
        d32 = dalloc(c, sizeof(float) * std::max<size_t>(1, v.size()));
        if (!d64 || !d32) return false;
        hold32.emplace_back(v.begin(), v.end());
        hold64.push_back(std::move(v));
        return hipMemcpyAsync(d64.get(), hold64.back().data(), sizeof(double) * hold64.back().size(),
                              hipMemcpyHostToDevice, s) == hipSuccess &&
               hipMemcpyAsync(d32.get(), hold32.back().data(), sizeof(float) * hold32.back().size(),
                              hipMemcpyHostToDevice, s) == hipSuccess;
    };
    namespace dn = gsx::dense;
    for (int k = 0; k < B.d; ++k) {
        const std::vector<double> &U = B.knots[k];
        const int p = B.p[k], nb = B.nb[k], lo = B.lo[k], hi = B.hi[k], n = hi - lo + 1;
        const bool dw = lo == 1, de = hi == nb - 2;
        if (n <= 0) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: no free functions in a direction");
        std::vector<double> gx(p + 1), gw(p + 1), K, M;
        gauss_rule(p + 1, gx.data(), gw.data());
        dn::bspline_1d_matrices(U, p, gx, gw, K, M);
        std::vector<double> Kf((size_t)n * n), Mf((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                Kf[(size_t)i * n + j] = beta * K[(size_t)(lo + i) * nb + (lo + j)];
                Mf[(size_t)i * n + j] = M[(size_t)(lo + i) * nb + (lo + j)];
            }
        // tilde space and complement (tildeSpaceBasis, :416-512); the
        // derivatives are scaled by powers of the largest knot span
        double hk = 0.0;
        for (size_t i = 0; i + 1 < U.size(); ++i)
            hk = std::max(hk, U[i + 1] - U[i]);
        std::vector<double> tL, cL, tR, cR;
        int nL, ncL, nR, ncR;
        if (!dn::tilde_oneside(U, p, true, dw, hk, tL, nL, cL, ncL) ||
            !dn::tilde_oneside(U, p, false, de, hk, tR, nR, cR, ncR))
            return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: boundary eigenproblem failed");
        const int mL = nL + ncL > 0 ? p - (dw ? 1 : 0) : 0, mR = nR + ncR > 0 ? p - (de ? 1 : 0) : 0;
        const int nI = n - mL - mR;
        int ct, cc;
        std::vector<double> Bt, Bl;   // n x ct, n x cc (row-major)
        if (nI <= 0) {
            ct = 0;
            cc = n;
            Bl.assign((size_t)n * n, 0.0);
            for (int i = 0; i < n; ++i) Bl[(size_t)i * n + i] = 1.0;
        } else {
            ct = nL + nI + nR;
            cc = ncL + ncR;
            Bt.assign((size_t)n * ct, 0.0);
            Bl.assign((size_t)n * std::max(cc, 1), 0.0);
            for (int i = 0; i < mL; ++i)
                for (int j = 0; j < nL; ++j) Bt[(size_t)i * ct + j] = tL[(size_t)i * nL + j];
            for (int j = 0; j < nI; ++j) Bt[(size_t)(mL + j) * ct + nL + j] = 1.0;
            for (int i = 0; i < mR; ++i)
                for (int j = 0; j < nR; ++j) Bt[(size_t)(mL + nI + i) * ct + nL + nI + j] = tR[(size_t)i * nR + j];
            for (int i = 0; i < mL; ++i)
                for (int j = 0; j < ncL; ++j) Bl[(size_t)i * cc + j] = cL[(size_t)i * ncL + j];
            for (int i = 0; i < mR; ++i)
                for (int j = 0; j < ncR; ++j) Bl[(size_t)(mL + nI + i) * cc + ncL + j] = cR[(size_t)i * ncR + j];
        }
        F->m[k] = n;
        F->ct[k] = ct;
        F->cc[k] = cc;
        auto gram = [&](const std::vector<double> &B, int cols, const std::vector<double> &A) {
            std::vector<double> G((size_t)cols * cols, 0.0), AB((size_t)n * cols, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < cols; ++j) {
                    double v = 0.0;
                    for (int l = 0; l < n; ++l) v += A[(size_t)i * n + l] * B[(size_t)l * cols + j];
                    AB[(size_t)i * cols + j] = v;
                }
            for (int a = 0; a < cols; ++a)
                for (int b = 0; b < cols; ++b) {
                    double v = 0.0;
                    for (int i = 0; i < n; ++i) v += B[(size_t)i * cols + a] * AB[(size_t)i * cols + b];
                    G[(size_t)a * cols + b] = v;
                }
            return G;
        };
        if (ct > 0) {
            // S = B~ (B~^T M B~)^-1 B~^T
            std::vector<double> Mt = gram(Bt, ct, Mf), X((size_t)ct * n);
            for (int a = 0; a < ct; ++a)
                for (int i = 0; i < n; ++i) X[(size_t)a * n + i] = Bt[(size_t)i * ct + a];
            if (!dn::lu_solve(ct, Mt, X, n)) return fail(c, GSX_ERR_ARG, "singular interior mass matrix");
            std::vector<double> S((size_t)n * n, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    double v = 0.0;
                    for (int a = 0; a < ct; ++a) v += Bt[(size_t)i * ct + a] * X[(size_t)a * n + j];
                    S[(size_t)i * n + j] = v;
                }
            if (!up(F->S64[k], F->S32[k], std::move(S))) return fail(c, GSX_ERR_OOM, "device allocation (smoother)");
        }
        if (cc > 0) {
            // B_compl = M^-1 B_l2compl; (K_c, M_c) = B_compl^T (K, M) B_compl, K_c Q = M_c Q diag(lambda)
            std::vector<double> Mcopy = Mf, Bc = Bl;
            Bc.resize((size_t)n * cc);
            if (!dn::lu_solve(n, Mcopy, Bc, cc)) return fail(c, GSX_ERR_ARG, "singular mass matrix");
            std::vector<double> Mc = gram(Bc, cc, Mf), Kc = gram(Bc, cc, Kf), lam, Q;
            if (!dn::gen_sym_eig(cc, Kc, Mc, lam, Q))
                return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: boundary eigenproblem failed");
            std::vector<double> W((size_t)n * cc, 0.0), WT((size_t)cc * n);
            for (int i = 0; i < n; ++i)
                for (int a = 0; a < cc; ++a) {
                    double v = 0.0;
                    for (int b = 0; b < cc; ++b) v += Bc[(size_t)i * cc + b] * Q[(size_t)b * cc + a];
                    W[(size_t)i * cc + a] = v;
                    WT[(size_t)a * n + i] = v;
                }
            DevPtr dummy;
            if (!up(F->W64[k], F->W32[k], std::move(W)) || !up(F->WT64[k], F->WT32[k], std::move(WT)) ||
                !up(F->lc[k], dummy, std::move(lam)))
                return fail(c, GSX_ERR_OOM, "device allocation (smoother)");
        }
    }
    F->size = (int64_t)F->m[0] * F->m[1] * F->m[2];
    for (int j = 0; j < 2; ++j) {
        F->t64[j] = dalloc(c, sizeof(double) * F->size);
        F->t32[j] = dalloc(c, sizeof(float) * F->size);
        if (!F->t64[j] || !F->t32[j]) return fail(c, GSX_ERR_OOM, "device allocation (smoother)");
    }
    F->xin64 = dalloc(c, sizeof(double) * F->size);
    if (!F->xin64) return fail(c, GSX_ERR_OOM, "device allocation (smoother)");
    HIP_OK(c, hipStreamSynchronize(s));
    out = std::move(F);
    return GSX_OK;
}

gsx_status gsx_scms_create(gsx_problem *P, double sigma, double alpha, double beta, gsx_fd **out)
{
    if (!P || !out) return GSX_ERR_ARG;
    gsx_ctx *c = P->ctx;
    ScmsBox B;
    if (!scms_box_of(P, B))
        return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother needs a single-patch problem whose free "
                                    "functions form a box");
    std::unique_ptr<gsx_fd> F;
    const gsx_status e = scms_build(c, B, sigma, alpha, beta, F);
    if (e != GSX_OK) return e;
    if (F->size != P->nfree) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: free box does not match");
    const HostPatch &hp = P->hp[0];
    const int pg = hp.p[0], pz = hp.d == 3 ? hp.p[0] : 0;
    for (int k = 0; k < 3; ++k) F->grid[k] = P->grid_dims[k];
    F->gs1 = P->grid_dims[0];
    F->gs2 = (int64_t)P->grid_dims[0] * P->grid_dims[1];
    F->goff = (P->box_lo[0] + pg) + F->gs1 * (P->box_lo[1] + pg) + F->gs2 * (P->box_lo[2] + pz);
    *out = F.release();
    return GSX_OK;
}

gsx_status gsx_fd_destroy(gsx_fd *F)
{
    delete F;
    return GSX_OK;
}

gsx_status gsx_fd_apply(gsx_fd *F, const double *r, double *z)
{
    if (!F || !r || !z) return GSX_ERR_ARG;
    gsx_ctx *c = F->ctx;
    KScope k(c, "fd_apply");
    if (F->kind == 1 && r == z) {
        HIP_OK(c, hipMemcpyAsync(F->xin64.get(), r, sizeof(double) * F->size, hipMemcpyDeviceToDevice, c->stream));
        r = dp<double>(F->xin64);
    }
    HIP_OK(c, fd_apply<double>(F, r, 0, F->m[0], (int64_t)F->m[0] * F->m[1], z, 0, F->m[0],
                               (int64_t)F->m[0] * F->m[1], nullptr, c->stream));
    return GSX_OK;
}

gsx_status gsx_cg_solve_fd(gsx_matrix *A, gsx_fd *F, int32_t precision, const double *b, double *x, double tol,
                           int32_t maxit, int32_t *iters, double *relres)
{
    if (!A || !F || !b || !x || maxit < 0) return GSX_ERR_ARG;
    gsx_ctx *c = A->ctx;
    if (F->ctx != c) return fail(c, GSX_ERR_ARG, "matrix and fast diagonalization live on different contexts");
    if (!A->st.available)
        return fail(c, GSX_ERR_ARG, "fast diagonalization needs the structured operator of a single-patch matrix");
    if (A->n != F->size || A->st.nx != F->grid[0] || A->st.ny != F->grid[1] || A->st.nz != F->grid[2])
        return fail(c, GSX_ERR_ARG, "matrix and fast diagonalization belong to different problems");
    if (precision == GSX_MIXED && !A->mixed)
        return fail(c, GSX_ERR_ARG, "mixed precision needs a matrix assembled with GSX_MIXED");
    const bool mixed = precision == GSX_MIXED;
    gsx_status e = cg_workspace(A, std::max(1, maxit), false, true);
    if (e != GSX_OK) return e;
    e = stencil_prepare(A, mixed);
    if (e != GSX_OK) return e;
    int it = 0;
    double rr = 0.0;
    e = cg_grid(A, GSX_PREC_NONE, 1.0, b, x, tol, maxit, mixed, &it, &rr, F);
    if (iters) *iters = it;
    if (relres) *relres = rr;
    return e;
}

gsx_status gsx_cg_history(gsx_matrix *A, double *hist, int32_t cap, int32_t *count)
{
    if (!A || !count) return GSX_ERR_ARG;
    *count = (int32_t)A->hist.size();
    if (hist)
        for (int i = 0; i < std::min<int>(cap, (int)A->hist.size()); ++i) hist[i] = A->hist[i];
    return GSX_OK;
}

gsx_status gsx_error_norms(gsx_problem *P, const double *sol, double *l2, double *h1semi)
{
    if (!P || !sol || (!l2 && !h1semi)) return GSX_ERR_ARG;
    gsx_ctx *c = P->ctx;
    if (P->geo) return fail(c, GSX_ERR_ARG, "error norms need the problem's own geometry");
    hipStream_t s = c->stream;
    const int64_t ne = std::max<int64_t>(1, P->nelem);
    DevPtr per = dalloc(c, sizeof(double) * 2 * (ne + 1));
    if (!per) return fail(c, GSX_ERR_OOM, "device allocation (error norms)");
    double *pl2 = dp<double>(per), *ph1 = pl2 + ne + 1;
    for (auto &g : P->groups)
        HIP_OK(c, gsx::launch_l2err(g.d, g.p, dp<PatchDev>(P->d_patches), dp<int>(g.d_gpatch),
                                    dp<int64_t>(g.d_gprefix), (int)g.patches.size(), g.nelem, c->gt,
                                    sol, P->nfree, pl2, h1semi ? ph1 : nullptr,
                                    P->d_fixed ? dp<double>(P->d_fixed) : nullptr, s));
    HIP_OK(c, gsx::launch_sum(pl2, P->nelem, pl2 + ne, s));
    if (h1semi) HIP_OK(c, gsx::launch_sum(ph1, P->nelem, ph1 + ne, s));
    double tot[2] = {0.0, 0.0};
    HIP_OK(c, hipMemcpyAsync(&tot[0], pl2 + ne, sizeof(double), hipMemcpyDeviceToHost, s));
    if (h1semi) HIP_OK(c, hipMemcpyAsync(&tot[1], ph1 + ne, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    if (l2) *l2 = std::sqrt(tot[0]);
    if (h1semi) *h1semi = std::sqrt(tot[1]);
    return GSX_OK;
}

gsx_status gsx_l2_error(gsx_problem *P, const double *sol, double *err)
{
    if (!err) return GSX_ERR_ARG;
    return gsx_error_norms(P, sol, err, nullptr);
}

}  // extern "C"

// ===========================================================================
// Geometric multigrid (gsMultiGridOp, src/gsMultiGrid/gsMultiGrid.hpp).
// Levels are built by coarsening every patch's knot vectors (dropping every
// other interior knot, the inverse of gsGridHierarchy::buildByRefinement,
// src/gsMultiGrid/gsGridHierarchy.hpp:27-101).  Transfers are the tensor
// products of 1D knot-insertion matrices restricted to free DOFs
// (gsMultiBasis::combineTransferMatrices, src/gsCore/gsMultiBasis.hpp:
// 137-178); coarse operators are Galerkin products P^T A P computed on the
// device; the coarsest level is solved exactly with a dense FP64 inverse.
// ===========================================================================
// Host copy of what the component decomposition of the subspace corrected
// mass smoother needs from a problem: patch sizes, knots and the DOF map.
struct PatchLayout {
    int d = 0, npatch = 0, nfree = 0;
    std::vector<HostPatch> hp;
    std::vector<std::vector<double>> knots;   // npatch * 3
    std::vector<int> dofs;
    std::vector<int64_t> boff;
};

static PatchLayout layout_of(const gsx_problem *P)
{
    PatchLayout L;
    L.d = P->d;
    L.npatch = P->npatch;
    L.nfree = P->nfree;
    L.hp = P->hp;
    L.knots = P->knots_h;
    L.dofs = P->dofs;
    L.boff = P->boff;
    return L;
}

// One component of the subspace corrected mass smoother: its free DOFs
// (global indices, component order) and local operator, the smoother of a
// box damped by tau, or the exact inverse of A restricted to the component.
struct ScmsComp {
    int n = 0, dim = 0;
    DevPtr idx;
    std::shared_ptr<gsx_fd> F;
    double tau = 1.0;
    DevPtr Ainv;   // n x n (direct components)
};
struct ScmsSplit {
    std::vector<ScmsComp> comps;
    int nmax = 0;
    DevPtr rc, zc;
};

struct MGLevel {
    gsx_problem *prob = nullptr;       // coarse levels own their problem
    int n = 0;
    int64_t nnz = 0;
    int lpr = 8;
    DevPtr rowptr, col, val, val32, diag, diag32;
    // transfer to the next finer level: P (n_fine x n), R = P^T (n x n_fine)
    int nfine = 0;
    int64_t pnnz = 0;
    DevPtr prp, pci, pv, pv32, rrp, rci, rv, rv32;
    DevPtr x, b, r, t;                 // FP64
    DevPtr x32, b32, r32, t32;         // FP32
    DevPtr d, d32;                     // polynomial smoother direction
    double lmax = 0.0;                 // Chebyshev: bound of lambda(D^-1 A)
    // Gauss-Seidel: rows by level set, forward and reverse sweep
    DevPtr gs_fwd, gs_rev;
    std::vector<int> gs_fwd_off, gs_rev_off;
    // subspace corrected mass smoother of the level's basis and its damping
    std::shared_ptr<gsx_fd> scms;
    // several components (patches, interfaces, vertices): section per component
    std::shared_ptr<ScmsSplit> split;
    double scms_tau = 0.0;
    // this level's smoother (levels >= 1): kind, damping of Jacobi/Richardson,
    // Chebyshev interval ratio
    int smoother = GSX_SMOOTH_JACOBI;
    double tau = 1.0, cheb_ratio = 30.0;
};

// Level sets of one Gauss-Seidel sweep over the CSR pattern: lev(i) = 1 +
// max lev(j) over the row's earlier columns j < i (forward) or later ones
// j > i (reverse); rows listed by level, ascending within a level.
static void gs_schedule(int n, const std::vector<int> &rp, const std::vector<int> &ci, bool reverse,
                        std::vector<int> &rows, std::vector<int> &off)
{
    std::vector<int> lev(n, 0);
    int maxl = 0;
    for (int k = 0; k < n; ++k) {
        const int i = reverse ? n - 1 - k : k;
        int l = 0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            const int c = ci[j];
            if (reverse ? c > i : c < i) l = std::max(l, lev[c] + 1);
        }
        lev[i] = l;
        maxl = std::max(maxl, l);
    }
    off.assign(maxl + 2, 0);
    for (int i = 0; i < n; ++i) off[lev[i] + 1]++;
    for (int l = 0; l <= maxl; ++l) off[l + 1] += off[l];
    rows.assign(n, 0);
    std::vector<int> at(off.begin(), off.end() - 1);
    for (int i = 0; i < n; ++i) rows[at[lev[i]]++] = i;
}

struct gsx_mg {
    gsx_ctx *ctx = nullptr;
    std::vector<MGLevel> L;            // 0 = coarsest
    int npre = 1, npost = 1, ncycles = 1;
    double tau = 1.0, cgc = 1.0;
    int precision = GSX_FP64;
    int npad = 0;
    // exact coarse solve by a banded Cholesky factor (band_b > 0) in reverse
    // Cuthill-McKee order, else by the dense inverse Dinv
    int band_b = 0;
    DevPtr band_pos, band_L, band_y;
    double scms_saved = 0.0;           // SCMS damping kept for finer levels
    std::vector<double> scms_saved_comp;   // the same per component (several components)
    PatchLayout fine_layout;           // finest level's patches and DOF map
    bool fine_layout_ok = false;
    DevPtr Dinv;
    DevPtr zin, zout;                  // finest FP64 staging for the mixed cycle
    ScmsBox fine_box;                  // free box of a single-patch finest level
    bool fine_box_ok = false;
    ~gsx_mg()
    {
        for (auto &l : L)
            if (l.prob) gsx_problem_destroy(l.prob);
    }
};

// 1D knot-insertion transfer coarse -> fine (fine knots = coarse knots plus
// the inserted ones), by single-knot Boehm insertion applied to the
// identity coefficients (gsBoehmRefine, src/gsNurbs/gsBoehm.hpp:161-226).
// Returns dense nf x nc (row-major).
static std::vector<double> transfer_1d(const std::vector<double> &coarse, const std::vector<double> &fine,
                                       int p, int &nf, int &nc)
{
    nc = (int)coarse.size() - p - 1;
    nf = (int)fine.size() - p - 1;
    std::vector<double> U = coarse;
    // rows of the transfer (fine function -> coarse coefficients), one
    // dense row per function; an insertion replaces the p - 1 rows
    // k-p+1..k-1 by p new rows and shifts the others
    std::vector<std::vector<double>> R((size_t)nc, std::vector<double>((size_t)nc, 0.0));
    for (int i = 0; i < nc; ++i) R[(size_t)i][(size_t)i] = 1.0;
    // knots of `fine` missing from `coarse` (multiset difference, in order)
    std::vector<double> ins;
    {
        size_t a = 0;
        for (size_t i = 0; i < fine.size(); ++i) {
            if (a < coarse.size() && coarse[a] == fine[i]) ++a;
            else ins.push_back(fine[i]);
        }
    }
    std::vector<std::vector<double>> mid;
    for (double u : ins) {
        int k = p;
        while (k + 1 < (int)U.size() && U[k + 1] <= u) ++k;
        mid.assign((size_t)p, std::vector<double>((size_t)nc));
        for (int i = k - p + 1; i <= k; ++i) {
            const double a = (u - U[i]) / (U[i + p] - U[i]);
            const std::vector<double> &r0 = R[(size_t)(i - 1)], &r1 = R[(size_t)i];
            std::vector<double> &dst = mid[(size_t)(i - (k - p + 1))];
            for (int c = 0; c < nc; ++c) dst[(size_t)c] = (1.0 - a) * r0[(size_t)c] + a * r1[(size_t)c];
        }
        std::vector<std::vector<double>> Rn;
        Rn.reserve(R.size() + 1);
        for (int i = 0; i <= k - p; ++i) Rn.push_back(std::move(R[(size_t)i]));
        for (auto &m : mid) Rn.push_back(std::move(m));
        for (size_t i = (size_t)k; i < R.size(); ++i) Rn.push_back(std::move(R[i]));
        R.swap(Rn);
        U.insert(U.begin() + k + 1, u);
    }
    std::vector<double> C((size_t)nf * nc);
    for (int i = 0; i < nf; ++i) std::memcpy(&C[(size_t)i * nc], R[(size_t)i].data(), sizeof(double) * nc);
    return C;
}

static gsx_status mg_level_vectors(gsx_ctx *c, MGLevel &l, bool mixed)
{
    const int n = std::max(1, l.n);
    l.x = dalloc(c, sizeof(double) * n);
    l.b = dalloc(c, sizeof(double) * n);
    l.r = dalloc(c, sizeof(double) * n);
    l.t = dalloc(c, sizeof(double) * n);
    if (!l.x || !l.b || !l.r || !l.t) return fail(c, GSX_ERR_OOM, "device allocation (multigrid vectors)");
    if (mixed) {
        l.x32 = dalloc(c, sizeof(float) * n);
        l.b32 = dalloc(c, sizeof(float) * n);
        l.r32 = dalloc(c, sizeof(float) * n);
        l.t32 = dalloc(c, sizeof(float) * n);
        if (!l.x32 || !l.b32 || !l.r32 || !l.t32) return fail(c, GSX_ERR_OOM, "device allocation (multigrid vectors)");
    }
    return GSX_OK;
}

static gsx_status scan_counts(gsx_ctx *c, const int *counts, int n, DevPtr &rowptr, int64_t &total)
{
    hipStream_t s = c->stream;
    rowptr = dalloc(c, sizeof(int) * (n + 1));
    DevPtr bsum = dalloc(c, sizeof(long long) * ((n + 1023) / 1024 + 2));
    DevPtr tot = dalloc(c, sizeof(long long));
    if (!rowptr || !bsum || !tot) return fail(c, GSX_ERR_OOM, "device allocation (scan)");
    HIP_OK(c, gsx::launch_scan(counts, n, dp<int>(rowptr), dp<long long>(bsum), dp<long long>(tot), s));
    long long t = 0;
    HIP_OK(c, hipMemcpyAsync(&t, tot.get(), sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    total = t;
    return GSX_OK;
}

// Reverse Cuthill-McKee order of a symmetric pattern (George and Liu,
// "Computer Solution of Large Sparse Positive Definite Systems", 1981):
// per connected part a breadth-first search from a node of least degree,
// restarted once from the last node that search reaches (a far node), with
// neighbours by ascending degree; the whole order reversed; ties go to the
// lower index.  pos[i] = position of row i; returns the half-bandwidth in
// that order.
static int rcm_order(int n, const std::vector<int> &rp, const std::vector<int> &ci, std::vector<int> &pos)
{
    std::vector<int> deg((size_t)n), order, level((size_t)n, -1);
    for (int i = 0; i < n; ++i) deg[(size_t)i] = rp[(size_t)i + 1] - rp[(size_t)i];
    std::vector<char> seen((size_t)n, 0);
    order.reserve((size_t)n);
    std::vector<int> nb;
    auto bfs = [&](int s, std::vector<int> &out) {
        // breadth-first over the unseen nodes from s; out in visiting order
        out.clear();
        std::vector<int> mark;
        out.push_back(s);
        level[(size_t)s] = 0;
        mark.push_back(s);
        for (size_t h = 0; h < out.size(); ++h) {
            const int v = out[h];
            nb.clear();
            for (int k = rp[(size_t)v]; k < rp[(size_t)v + 1]; ++k) {
                const int u = ci[(size_t)k];
                if (u != v && !seen[(size_t)u] && level[(size_t)u] < 0) {
                    level[(size_t)u] = level[(size_t)v] + 1;
                    mark.push_back(u);
                    nb.push_back(u);
                }
            }
            std::sort(nb.begin(), nb.end(), [&](int a, int b) {
                return deg[(size_t)a] != deg[(size_t)b] ? deg[(size_t)a] < deg[(size_t)b] : a < b;
            });
            out.insert(out.end(), nb.begin(), nb.end());
        }
        for (int m : mark) level[(size_t)m] = -1;
    };
    std::vector<int> part;
    for (;;) {
        int s = -1;
        for (int i = 0; i < n; ++i)
            if (!seen[(size_t)i] && (s < 0 || deg[(size_t)i] < deg[(size_t)s])) s = i;
        if (s < 0) break;
        bfs(s, part);
        bfs(part.back(), part);
        for (int v : part) {
            seen[(size_t)v] = 1;
            order.push_back(v);
        }
    }
    std::reverse(order.begin(), order.end());
    pos.assign((size_t)n, 0);
    for (int k = 0; k < n; ++k) pos[(size_t)order[(size_t)k]] = k;
    int b = 0;
    for (int i = 0; i < n; ++i)
        for (int k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k) b = std::max(b, std::abs(pos[(size_t)i] - pos[(size_t)ci[(size_t)k]]));
    return b;
}

// Per-level diagonals, FP32 copies and vectors, then the exact coarse solve
// (the dense inverse of the coarsest operator or, above the context's
// coarse_dense_max rows, its banded FP64 Cholesky factor).
static gsx_status mg_finish(gsx_mg *mg)
{
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    // per-level diagonals, FP32 copies, vectors
    const int nlevels = (int)mg->L.size();
    const bool mixed = mg->precision == GSX_MIXED;
    for (int l = 0; l < nlevels; ++l) {
        mg->L[l].tau = mg->tau;
        MGLevel &V = mg->L[l];
        V.diag = dalloc(c, sizeof(double) * std::max(1, V.n));
        if (!V.diag) return fail(c, GSX_ERR_OOM, "device allocation (diag)");
        HIP_OK(c, gsx::launch_diag_f64(V.n, dp<int>(V.rowptr), dp<int>(V.col), dp<double>(V.val), dp<double>(V.diag), s));
        if (mixed) {
            V.val32 = dalloc(c, sizeof(float) * std::max<int64_t>(1, V.nnz));
            V.diag32 = dalloc(c, sizeof(float) * std::max(1, V.n));
            if (!V.val32 || !V.diag32) return fail(c, GSX_ERR_OOM, "device allocation (fp32 level)");
            HIP_OK(c, gsx::launch_conv_f64_f32(V.nnz, dp<double>(V.val), dp<float>(V.val32), s));
            HIP_OK(c, gsx::launch_conv_f64_f32(V.n, dp<double>(V.diag), dp<float>(V.diag32), s));
            if (l < nlevels - 1) {
                V.pv32 = dalloc(c, sizeof(float) * std::max<int64_t>(1, V.pnnz));
                V.rv32 = dalloc(c, sizeof(float) * std::max<int64_t>(1, V.pnnz));
                if (!V.pv32 || !V.rv32) return fail(c, GSX_ERR_OOM, "device allocation (fp32 transfer)");
                HIP_OK(c, gsx::launch_conv_f64_f32(V.pnnz, dp<double>(V.pv), dp<float>(V.pv32), s));
                HIP_OK(c, gsx::launch_conv_f64_f32(V.pnnz, dp<double>(V.rv), dp<float>(V.rv32), s));
            }
        }
        GSX_TRY(mg_level_vectors(c, V, mixed));
    }
    // exact coarse solve: the dense inverse up to coarse_dense_max rows (one
    // GEMV per solve); above, a banded FP64 Cholesky factor in reverse
    // Cuthill-McKee order (its two triangular solves run panel after panel,
    // slower than the GEMV where both fit)
    {
        KScope ki(c, "mg_coarse_inverse");
        MGLevel &C0 = mg->L[0];
        const int n0 = C0.n;
        if ((int64_t)n0 > c->coarse_dense_max) {
            std::vector<int> rp((size_t)n0 + 1), ci((size_t)C0.nnz), pos;
            HIP_OK(c, hipMemcpyAsync(rp.data(), C0.rowptr.get(), sizeof(int) * (n0 + 1), hipMemcpyDeviceToHost, s));
            HIP_OK(c, hipMemcpyAsync(ci.data(), C0.col.get(), sizeof(int) * C0.nnz, hipMemcpyDeviceToHost, s));
            HIP_OK(c, hipStreamSynchronize(s));
            const int b = rcm_order(n0, rp, ci, pos);
            if (b <= 4096) {
                mg->band_b = b;
                mg->band_pos = dalloc(c, sizeof(int) * n0);
                mg->band_L = dalloc(c, sizeof(double) * (size_t)n0 * (b + 1));
                mg->band_y = dalloc(c, sizeof(double) * n0);
                DevPtr errf = dalloc(c, sizeof(int));
                if (!mg->band_pos || !mg->band_L || !mg->band_y || !errf)
                    return fail(c, GSX_ERR_OOM, "device allocation (coarse factor)");
                HIP_OK(c, hipMemcpyAsync(mg->band_pos.get(), pos.data(), sizeof(int) * n0, hipMemcpyHostToDevice, s));
                HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int), s));
                HIP_OK(c, gsx::launch_band_factor(n0, b, dp<int>(C0.rowptr), dp<int>(C0.col), dp<double>(C0.val),
                                                  dp<int>(mg->band_pos), dp<double>(mg->band_L), dp<int>(errf), s));
                int errh = 0;
                HIP_OK(c, hipMemcpyAsync(&errh, errf.get(), sizeof(int), hipMemcpyDeviceToHost, s));
                HIP_OK(c, hipStreamSynchronize(s));
                if (errh) return fail(c, GSX_ERR_ARG, "coarsest operator is not positive definite (Cholesky pivot)");
                return GSX_OK;
            }
        }
        if ((int64_t)n0 > 32768)
            return fail(c, GSX_ERR_ARG, "coarsest level too large for the dense inverse and its band too wide");
        mg->npad = gsx::dense_inverse_npad(n0);
        mg->Dinv = dalloc(c, sizeof(double) * (size_t)mg->npad * mg->npad);
        DevPtr W = dalloc(c, sizeof(double) * (size_t)gsx::dense_inverse_work_doubles(mg->npad));
        if (!mg->Dinv || !W) return fail(c, GSX_ERR_OOM, "device allocation (coarse inverse)");
        HIP_OK(c, gsx::launch_dense_inverse(n0, mg->npad, dp<int>(C0.rowptr), dp<int>(C0.col), dp<double>(C0.val),
                                            dp<double>(mg->Dinv), dp<double>(W), s));
        HIP_OK(c, hipStreamSynchronize(s));
    }
    return GSX_OK;
}

// 1D knot-insertion transfers, computed once per distinct (coarse knots,
// fine knots, p)
struct TransferCache {
    struct Item { std::vector<double> a, b; int p, nf, nc; std::vector<double> T; };
    std::deque<Item> items;   // references stay valid as items are added
    const std::vector<double> &get(const std::vector<double> &a, const std::vector<double> &b, int p, int &nf,
                                   int &nc)
    {
        for (const Item &it : items)
            if (it.p == p && it.a == a && it.b == b) {
                nf = it.nf;
                nc = it.nc;
                return it.T;
            }
        Item it{a, b, p, 0, 0, {}};
        it.T = transfer_1d(a, b, p, it.nf, it.nc);
        nf = it.nf;
        nc = it.nc;
        items.push_back(std::move(it));
        return items.back().T;
    }
};

// P: the knot-insertion transfer from the basis of cp to that of fp (same
// degrees, cp's knots a subset of fp's) restricted to free DOFs, rows = fp's
// nf free DOFs (gsMultiBasis::combineTransferMatrices, src/gsCore/
// gsMultiBasis.hpp:137-178), into C.prp/pci/pv.  With tblk, also the per
// (patch, direction, fine element) 1D blocks of the element Galerkin path
// at toff.
static gsx_status mg_knot_transfer(gsx_ctx *c, gsx_problem *fp, gsx_problem *cp, int nf, MGLevel &C,
                                   TransferCache &tc, std::vector<double> *tblk, std::vector<int64_t> *toff)
{
    hipStream_t s = c->stream;
    const int npatch = fp->npatch, d = fp->d;
    // 1D transfers per (patch, direction)
    struct T1Dh { int rp_off, ent_off, nf, nc; };
    std::vector<T1Dh> t1d((size_t)npatch * 3, T1Dh{0, 0, 1, 1});
    std::vector<int> trp, tcol;
    std::vector<double> tval;
    std::vector<int> fnb((size_t)npatch * 3, 1), cnb((size_t)npatch * 3, 1), cdoff(npatch, 0);
    // element path: per (patch, direction, fine span e) the block
    // T[m][a] = Tm[e + m][e / 2 + a], (p+1)^2 doubles
    if (toff) toff->assign((size_t)npatch * 3, 0);
    for (int k = 0; k < npatch; ++k) {
        cdoff[k] = (int)cp->boff[k];
        for (int j = 0; j < 3; ++j) {
            fnb[3 * k + j] = fp->hp[k].nb[j];
            cnb[3 * k + j] = cp->hp[k].nb[j];
        }
        for (int j = 0; j < 3; ++j) {
            T1Dh &T = t1d[3 * k + j];
            T.rp_off = (int)trp.size();
            T.ent_off = (int)tcol.size();
            if (j >= d) {
                trp.push_back(0);
                trp.push_back(1);
                tcol.push_back(0);
                tval.push_back(1.0);
                continue;
            }
            int nf1 = 0, nc1 = 0;   // 1D fine and coarse function counts
            const std::vector<double> &Tm = tc.get(cp->knots_h[(size_t)3 * k + j],
                                                   fp->knots_h[(size_t)3 * k + j], fp->hp[k].p[j], nf1, nc1);
            T.nf = nf1;
            T.nc = nc1;
            if (tblk) {
                const int p1 = fp->hp[k].p[j] + 1, nelf = fp->hp[k].nel[j];
                (*toff)[(size_t)3 * k + j] = (int64_t)tblk->size();
                for (int e = 0; e < nelf; ++e)
                    for (int m = 0; m < p1; ++m)
                        for (int a = 0; a < p1; ++a) tblk->push_back(Tm[(size_t)(e + m) * nc1 + e / 2 + a]);
            }
            int cnt = 0;
            for (int i = 0; i < nf1; ++i) {
                trp.push_back(cnt);
                for (int cc = 0; cc < nc1; ++cc)
                    if (Tm[(size_t)i * nc1 + cc] != 0.0) {
                        tcol.push_back(cc);
                        tval.push_back(Tm[(size_t)i * nc1 + cc]);
                        ++cnt;
                    }
            }
            trp.push_back(cnt);
        }
    }
    // rp entries are relative to ent_off
    DevPtr d_t1d = dalloc(c, sizeof(T1Dh) * t1d.size());
    DevPtr d_trp = dalloc(c, sizeof(int) * trp.size()), d_tcol = dalloc(c, sizeof(int) * tcol.size());
    DevPtr d_tval = dalloc(c, sizeof(double) * tval.size());
    DevPtr d_fnb = dalloc(c, sizeof(int) * fnb.size()), d_cnb = dalloc(c, sizeof(int) * cnb.size());
    DevPtr d_cdoff = dalloc(c, sizeof(int) * npatch);
    DevPtr counts = dalloc(c, sizeof(int) * (nf + 1)), errf = dalloc(c, sizeof(int) * 2);
    if (!d_t1d || !d_trp || !d_tcol || !d_tval || !d_fnb || !d_cnb || !d_cdoff || !counts || !errf)
        return fail(c, GSX_ERR_OOM, "device allocation (transfer)");
    HIP_OK(c, hipMemcpyAsync(d_t1d.get(), t1d.data(), sizeof(T1Dh) * t1d.size(), hipMemcpyHostToDevice, s));
    HIP_OK(c, hipMemcpyAsync(d_trp.get(), trp.data(), sizeof(int) * trp.size(), hipMemcpyHostToDevice, s));
    HIP_OK(c, hipMemcpyAsync(d_tcol.get(), tcol.data(), sizeof(int) * tcol.size(), hipMemcpyHostToDevice, s));
    HIP_OK(c, hipMemcpyAsync(d_tval.get(), tval.data(), sizeof(double) * tval.size(), hipMemcpyHostToDevice, s));
    HIP_OK(c, hipMemcpyAsync(d_fnb.get(), fnb.data(), sizeof(int) * fnb.size(), hipMemcpyHostToDevice, s));
    HIP_OK(c, hipMemcpyAsync(d_cnb.get(), cnb.data(), sizeof(int) * cnb.size(), hipMemcpyHostToDevice, s));
    HIP_OK(c, hipMemcpyAsync(d_cdoff.get(), cdoff.data(), sizeof(int) * npatch, hipMemcpyHostToDevice, s));
    HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int) * 2, s));
    HIP_OK(c, gsx::launch_transfer(0, nf, dp<int>(fp->d_occ_ptr), dp<int>(fp->d_occ_patch),
                                   dp<int64_t>(fp->d_occ_local), dp<int>(d_fnb), d_t1d.get(), dp<int>(d_trp),
                                   dp<int>(d_tcol), dp<double>(d_tval), dp<int>(d_cdoff), dp<int>(cp->d_dofs),
                                   dp<int>(d_cnb), d, C.n, dp<int>(counts), dp<int>(errf), nullptr, nullptr,
                                   nullptr, s));
    int64_t pn = 0;
    GSX_TRY(scan_counts(c, dp<int>(counts), nf, C.prp, pn));
    C.nfine = nf;
    C.pnnz = pn;
    C.pci = dalloc(c, sizeof(int) * std::max<int64_t>(1, pn));
    C.pv = dalloc(c, sizeof(double) * std::max<int64_t>(1, pn));
    if (!C.pci || !C.pv) return fail(c, GSX_ERR_OOM, "device allocation (transfer)");
    HIP_OK(c, gsx::launch_transfer(1, nf, dp<int>(fp->d_occ_ptr), dp<int>(fp->d_occ_patch),
                                   dp<int64_t>(fp->d_occ_local), dp<int>(d_fnb), d_t1d.get(), dp<int>(d_trp),
                                   dp<int>(d_tcol), dp<double>(d_tval), dp<int>(d_cdoff), dp<int>(cp->d_dofs),
                                   dp<int>(d_cnb), d, C.n, nullptr, nullptr, dp<int>(C.prp), dp<int>(C.pci),
                                   dp<double>(C.pv), s));
    int errh[2] = {0, 0};
    HIP_OK(c, hipMemcpyAsync(errh, errf.get(), sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    if (errh[0] || errh[1]) return fail(c, GSX_ERR_ARG, "multigrid transfer: row buffer overflow");
    return GSX_OK;
}

// R = P^T (rows sorted by column) into C.rrp/rci/rv; C.n coarse rows, nf fine.
static gsx_status mg_transpose_transfer(gsx_ctx *c, MGLevel &C, int nf)
{
    hipStream_t s = c->stream;
    DevPtr tcount = dalloc(c, sizeof(int) * (C.n + 1)), cursor = dalloc(c, sizeof(int) * (C.n + 1));
    if (!tcount || !cursor) return fail(c, GSX_ERR_OOM, "device allocation (restriction)");
    HIP_OK(c, hipMemsetAsync(tcount.get(), 0, sizeof(int) * (C.n + 1), s));
    HIP_OK(c, hipMemsetAsync(cursor.get(), 0, sizeof(int) * (C.n + 1), s));
    HIP_OK(c, gsx::launch_transpose(0, nf, dp<int>(C.prp), dp<int>(C.pci), nullptr, dp<int>(tcount), nullptr,
                                    nullptr, nullptr, nullptr, 0, s));
    int64_t rn = 0;
    GSX_TRY(scan_counts(c, dp<int>(tcount), C.n, C.rrp, rn));
    C.rci = dalloc(c, sizeof(int) * std::max<int64_t>(1, rn));
    C.rv = dalloc(c, sizeof(double) * std::max<int64_t>(1, rn));
    if (!C.rci || !C.rv) return fail(c, GSX_ERR_OOM, "device allocation (restriction)");
    HIP_OK(c, gsx::launch_transpose(1, nf, dp<int>(C.prp), dp<int>(C.pci), dp<double>(C.pv), nullptr,
                                    dp<int>(C.rrp), dp<int>(cursor), dp<int>(C.rci), dp<double>(C.rv), 0, s));
    HIP_OK(c, gsx::launch_transpose(2, 0, nullptr, nullptr, nullptr, nullptr, dp<int>(C.rrp), nullptr,
                                    dp<int>(C.rci), dp<double>(C.rv), C.n, s));
    return GSX_OK;
}

// Galerkin operator A_c = R (A_f P) (src/gsMultiGrid/gsMultiGrid.hpp:70-104)
// into the coarse pattern C.rowptr/col, values C.val: B = A_f P row by row
// in ELL form, then R B; overflow codes in err[0] (B) and err[1] (pattern).
static gsx_status mg_galerkin_sparse(gsx_ctx *c, MGLevel &C, const MGLevel &F, int d, int *err)
{
    hipStream_t s = c->stream;
    const int WB = gsx::ap_width(d);
    DevPtr bcnt = dalloc(c, sizeof(int) * std::max(1, F.n)), bci = dalloc(c, sizeof(int) * (size_t)std::max(1, F.n) * WB);
    DevPtr bv = dalloc(c, sizeof(double) * (size_t)std::max(1, F.n) * WB);
    if (!bcnt || !bci || !bv) return fail(c, GSX_ERR_OOM, "device allocation (Galerkin product)");
    HIP_OK(c, gsx::launch_ap(d, F.n, dp<int>(F.rowptr), dp<int>(F.col), dp<double>(F.val), dp<int>(C.prp),
                             dp<int>(C.pci), dp<double>(C.pv), dp<int>(bcnt), dp<int>(bci), dp<double>(bv), err, s));
    HIP_OK(c, gsx::launch_rap(d, C.n, dp<int>(C.rrp), dp<int>(C.rci), dp<double>(C.rv), dp<int>(bcnt), dp<int>(bci),
                              dp<double>(bv), dp<int>(C.rowptr), dp<int>(C.col), dp<double>(C.val), err + 1, s));
    return GSX_OK;
}

extern "C" gsx_status gsx_mg_create(gsx_problem *fine, gsx_matrix *A, int32_t nlevels, int32_t npre,
                                    int32_t npost, int32_t ncycles, double damping, double cgc_damping,
                                    int32_t precision, gsx_mg **out)
{
    if (!fine || !A || !out || nlevels < 1 || npre < 0 || npost < 0 || ncycles < 1) return GSX_ERR_ARG;
    gsx_ctx *c = fine->ctx;
    if (ncycles != 1) return fail(c, GSX_ERR_ARG, "only V-cycles (NumCycles = 1) are implemented");
    if (A->n != fine->nfree) return fail(c, GSX_ERR_ARG, "matrix does not belong to the problem");
    if (precision != GSX_FP64 && precision != GSX_MIXED) return fail(c, GSX_ERR_ARG, "unknown precision");
    hipStream_t s = c->stream;
    KScope ks(c, "mg_setup");
    auto mg = std::make_unique<gsx_mg>();
    mg->ctx = c;
    mg->npre = npre;
    mg->npost = npost;
    mg->ncycles = ncycles;
    mg->tau = damping;
    mg->cgc = cgc_damping;
    mg->precision = precision;
    mg->L.resize(nlevels);
    mg->fine_box_ok = scms_box_of(fine, mg->fine_box);
    mg->fine_layout = layout_of(fine);
    mg->fine_layout_ok = true;
    // finest level = the given matrix
    {
        MGLevel &f = mg->L[nlevels - 1];
        f.n = A->n;
        f.nnz = A->nnz;
        f.lpr = A->lpr;
        f.rowptr = A->rowptr;
        f.col = A->col;
        f.val = A->val;
    }
    // knot vectors of all levels
    const int npatch = fine->npatch, d = fine->d;
    std::vector<std::vector<std::vector<double>>> knots(nlevels);
    knots[nlevels - 1] = fine->knots_h;
    for (int l = nlevels - 2; l >= 0; --l) {
        knots[l].resize((size_t)npatch * 3);
        for (int k = 0; k < npatch; ++k)
            for (int j = 0; j < d; ++j) {
                const std::vector<double> &U = knots[l + 1][3 * k + j];
                const int p = fine->hp[k].p[j];
                const int nel = (int)U.size() - 2 * p - 1;
                if (nel % 2 != 0 || nel < 2)
                    return fail(c, GSX_ERR_ARG, "multigrid needs an even number of elements per direction on every level");
                std::vector<double> V(U.begin(), U.begin() + p + 1);
                for (int e = 2; e < nel; e += 2) V.push_back(U[p + e]);
                V.insert(V.end(), U.end() - p - 1, U.end());
                knots[l][3 * k + j] = V;
            }
    }
    // coarse problems (geometry is not used by the Galerkin levels)
    for (int l = nlevels - 2; l >= 0; --l) {
        KScope kl(c, "mg_levels");
        std::vector<gsx_patch_desc> desc(npatch);
        std::vector<std::vector<double>> ctrl(npatch);
        for (int k = 0; k < npatch; ++k) {
            gsx_patch_desc &D = desc[k];
            std::memset(&D, 0, sizeof(D));
            D.d = d;
            int64_t nbas = 1;
            for (int j = 0; j < d; ++j) {
                D.p[j] = fine->hp[k].p[j];
                D.nknots[j] = (int)knots[l][3 * k + j].size();
                D.knots_host[j] = knots[l][3 * k + j].data();
                nbas *= D.nknots[j] - D.p[j] - 1;
            }
            ctrl[k].assign((size_t)nbas * d, 0.0);
            D.ctrl_host = ctrl[k].data();
            D.fkind = 1;
        }
        gsx_problem *cp = nullptr;
        gsx_status st = gsx_problem_create(c, npatch, desc.data(), (int)fine->dir_sides_h.size() / 2,
                                           fine->dir_sides_h.empty() ? nullptr : fine->dir_sides_h.data(),
                                           (int)fine->ifaces_h.size() / 8,
                                           fine->ifaces_h.empty() ? nullptr : fine->ifaces_h.data(), &cp);
        if (st != GSX_OK) return st;
        mg->L[l].prob = cp;
        mg->L[l].n = cp->nfree;
        st = pattern_build(cp);
        if (st != GSX_OK) return st;
    }
    // Galerkin levels from element matrices when A is the assembled matrix of
    // `fine` and the element matrices of two neighbouring levels fit the
    // element budget: the fine element matrices are recomputed, then each
    // coarse level's come from its children's (k_coarsen) and are gathered
    // into the coarse pattern.  Otherwise the global product P^T (A P).
    bool elem_path = A->src == fine && nlevels > 1;
    for (const Group &g : fine->groups) elem_path = elem_path && gsx::coarsen_supported(g.d, g.p);
    if (elem_path) {
        const int64_t kc = nlevels > 1 ? mg->L[nlevels - 2].prob->kdoubles : 0;
        elem_path = (fine->kdoubles + fine->fdoubles + kc) * (int64_t)sizeof(double) <= c->elem_budget;
    }
    DevPtr Kf;
    if (elem_path && fine->kept_K && fine->kept_serial == A->src_serial) {
        Kf = std::move(fine->kept_K);   // this matrix's own element matrices
    } else if (elem_path) {
        KScope ke(c, "mg_elem");
        Kf = dalloc(c, sizeof(double) * std::max<int64_t>(1, fine->kdoubles));
        DevPtr Ff = dalloc(c, sizeof(double) * std::max<int64_t>(1, fine->fdoubles));
        if (!Kf || !Ff) return fail(c, GSX_ERR_OOM, "device allocation (element matrices)");
        for (const Group &g : fine->groups)
            HIP_OK(c, gsx::launch_elem(g.d, g.p, dp<PatchDev>(fine->d_patches), dp<int>(g.d_gpatch),
                                       dp<int64_t>(g.d_gprefix), (int)g.patches.size(), g.nelem, c->gt,
                                       dp<double>(Kf), dp<double>(Ff), dp<double>(fine->d_detmin), s));
    }
    TransferCache tcache;
    // transfers and Galerkin products, finest first
    for (int l = nlevels - 2; l >= 0; --l) {
        MGLevel &C = mg->L[l], &F = mg->L[l + 1];
        KScope kt(c, "mg_transfer_build");
        gsx_problem *fp = (l + 1 == nlevels - 1) ? fine : F.prob;
        gsx_problem *cp = C.prob;
        std::vector<double> tblk;
        std::vector<int64_t> toff;
        GSX_TRY(mg_knot_transfer(c, fp, cp, F.n, C, tcache, elem_path ? &tblk : nullptr, &toff));
        GSX_TRY(mg_transpose_transfer(c, C, F.n));
        DevPtr errf = dalloc(c, sizeof(int) * 2);
        if (!errf) return fail(c, GSX_ERR_OOM, "device allocation (transfer)");
        HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int) * 2, s));
        C.rowptr = cp->d_rowptr;
        C.col = cp->d_col;
        C.nnz = cp->nnz;
        C.lpr = gsx::spmv_lpr(C.n ? (double)C.nnz / C.n : 1.0);
        C.val = dalloc(c, sizeof(double) * std::max<int64_t>(1, C.nnz));
        if (!C.val) return fail(c, GSX_ERR_OOM, "device allocation (Galerkin product)");
        if (elem_path) {
            // coarse element matrices from the children's, then the coarse gather
            KScope kg(c, "mg_galerkin");
            DevPtr Kc = dalloc(c, sizeof(double) * std::max<int64_t>(1, cp->kdoubles));
            DevPtr d_tb = dalloc(c, sizeof(double) * std::max<size_t>(1, tblk.size()));
            DevPtr d_toff = dalloc(c, sizeof(int64_t) * toff.size());
            if (!Kc || !d_tb || !d_toff) return fail(c, GSX_ERR_OOM, "device allocation (coarse element matrices)");
            HIP_OK(c, hipMemcpyAsync(d_tb.get(), tblk.data(), sizeof(double) * tblk.size(), hipMemcpyHostToDevice, s));
            HIP_OK(c, hipMemcpyAsync(d_toff.get(), toff.data(), sizeof(int64_t) * toff.size(), hipMemcpyHostToDevice, s));
            for (const Group &g : cp->groups)
                HIP_OK(c, gsx::launch_coarsen(g.d, g.p, dp<PatchDev>(cp->d_patches), dp<PatchDev>(fp->d_patches),
                                              dp<int>(g.d_gpatch), dp<int64_t>(g.d_gprefix), (int)g.patches.size(),
                                              g.nelem, dp<double>(d_tb), dp<int64_t>(d_toff), dp<double>(Kf),
                                              dp<double>(Kc), s));
            HIP_OK(c, gsx::launch_gather(dp<PatchDev>(cp->d_patches), dp<int>(cp->d_occ_ptr), dp<int>(cp->d_occ_patch),
                                         dp<int64_t>(cp->d_occ_local), C.n, dp<int>(C.rowptr), dp<int>(C.col),
                                         dp<double>(Kc), nullptr, dp<double>(C.val), nullptr, nullptr, cp->maxrow,
                                         nullptr, s));
            Kf = std::move(Kc);
        } else {
            GSX_TRY(mg_galerkin_sparse(c, C, F, d, dp<int>(errf)));
        }
        int errh[2] = {0, 0};
        HIP_OK(c, hipMemcpyAsync(errh, errf.get(), sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipStreamSynchronize(s));
        if (errh[0] || errh[1]) {
            char buf[160];
            std::snprintf(buf, sizeof(buf), "multigrid level %d: transfer/Galerkin buffer overflow (codes %d, %d)", l,
                          errh[0], errh[1]);
            return fail(c, GSX_ERR_ARG, buf);
        }
    }
    GSX_TRY(mg_finish(mg.get()));
    *out = mg.release();
    return GSX_OK;
}

// gsMultiGridOp::make(fineMatrix, transferMatrices) + init
// (src/gsMultiGrid/gsMultiGrid.hpp:19-35,70-94): the caller's transfers
// (transfer l prolongates level l to level l + 1, n_{l+1} x n_l row-major
// CSR, level nlevels - 1 = A), coarse operators P^T A P formed on the device
// (B = A P, then the pattern and values of P^T B), restriction P^T.
extern "C" gsx_status gsx_mg_create_transfers(gsx_matrix *A, int32_t ntransfer, const int32_t *ncoarse,
                                              const int32_t *const *prow, const int32_t *const *pcol,
                                              const double *const *pval, int32_t npre, int32_t npost,
                                              int32_t ncycles, double damping, double cgc_damping,
                                              int32_t precision, gsx_mg **out)
{
    if (!A || !out || ntransfer < 0 || npre < 0 || npost < 0 || ncycles < 1) return GSX_ERR_ARG;
    if (ntransfer > 0 && (!ncoarse || !prow || !pcol || !pval)) return GSX_ERR_ARG;
    gsx_ctx *c = A->ctx;
    if (ncycles != 1) return fail(c, GSX_ERR_ARG, "only V-cycles (NumCycles = 1) are implemented");
    if (precision != GSX_FP64 && precision != GSX_MIXED) return fail(c, GSX_ERR_ARG, "unknown precision");
    hipStream_t s = c->stream;
    KScope ks(c, "mg_setup");
    const int nlevels = ntransfer + 1;
    auto mg = std::make_unique<gsx_mg>();
    mg->ctx = c;
    mg->npre = npre;
    mg->npost = npost;
    mg->ncycles = ncycles;
    mg->tau = damping;
    mg->cgc = cgc_damping;
    mg->precision = precision;
    mg->L.resize(nlevels);
    mg->fine_box_ok = A->src && scms_box_of(A->src, mg->fine_box);
    {
        MGLevel &f = mg->L[nlevels - 1];
        f.n = A->n;
        f.nnz = A->nnz;
        f.lpr = A->lpr;
        f.rowptr = A->rowptr;
        f.col = A->col;
        f.val = A->val;
    }
    for (int l = nlevels - 2; l >= 0; --l) {
        MGLevel &C = mg->L[l], &F = mg->L[l + 1];
        const int nf = F.n, nc = ncoarse[l];
        const int32_t *rp = prow[l], *ci = pcol[l];
        if (nc < 0 || !rp || !ci || !pval[l]) return fail(c, GSX_ERR_ARG, "transfer matrix missing");
        if (rp[0] != 0) return fail(c, GSX_ERR_ARG, "transfer matrix: rowptr[0] != 0");
        for (int i = 0; i < nf; ++i) {
            if (rp[i + 1] < rp[i]) return fail(c, GSX_ERR_ARG, "transfer matrix: rowptr not ascending");
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] < 0 || ci[k] >= nc || (k > rp[i] && ci[k] <= ci[k - 1])) {
                    char buf[160];
                    std::snprintf(buf, sizeof(buf),
                                  "transfer matrix %d: row %d has a column out of range or not ascending", l, i);
                    return fail(c, GSX_ERR_ARG, buf);
                }
        }
        const int64_t pn = rp[nf];
        C.n = nc;
        C.nfine = nf;
        C.pnnz = pn;
        C.prp = dalloc(c, sizeof(int) * (nf + 1));
        C.pci = dalloc(c, sizeof(int) * std::max<int64_t>(1, pn));
        C.pv = dalloc(c, sizeof(double) * std::max<int64_t>(1, pn));
        if (!C.prp || !C.pci || !C.pv) return fail(c, GSX_ERR_OOM, "device allocation (transfer)");
        HIP_OK(c, hipMemcpyAsync(C.prp.get(), rp, sizeof(int) * (nf + 1), hipMemcpyHostToDevice, s));
        HIP_OK(c, hipMemcpyAsync(C.pci.get(), ci, sizeof(int) * pn, hipMemcpyHostToDevice, s));
        HIP_OK(c, hipMemcpyAsync(C.pv.get(), pval[l], sizeof(double) * pn, hipMemcpyHostToDevice, s));
        // R = P^T
        DevPtr tcount = dalloc(c, sizeof(int) * (nc + 1)), cursor = dalloc(c, sizeof(int) * (nc + 1));
        DevPtr errf = dalloc(c, sizeof(int) * 2);
        if (!tcount || !cursor || !errf) return fail(c, GSX_ERR_OOM, "device allocation (restriction)");
        HIP_OK(c, hipMemsetAsync(tcount.get(), 0, sizeof(int) * (nc + 1), s));
        HIP_OK(c, hipMemsetAsync(cursor.get(), 0, sizeof(int) * (nc + 1), s));
        HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int) * 2, s));
        HIP_OK(c, gsx::launch_transpose(0, nf, dp<int>(C.prp), dp<int>(C.pci), nullptr, dp<int>(tcount), nullptr,
                                        nullptr, nullptr, nullptr, 0, s));
        int64_t rn = 0;
        GSX_TRY(scan_counts(c, dp<int>(tcount), nc, C.rrp, rn));
        C.rci = dalloc(c, sizeof(int) * std::max<int64_t>(1, rn));
        C.rv = dalloc(c, sizeof(double) * std::max<int64_t>(1, rn));
        if (!C.rci || !C.rv) return fail(c, GSX_ERR_OOM, "device allocation (restriction)");
        HIP_OK(c, gsx::launch_transpose(1, nf, dp<int>(C.prp), dp<int>(C.pci), dp<double>(C.pv), nullptr,
                                        dp<int>(C.rrp), dp<int>(cursor), dp<int>(C.rci), dp<double>(C.rv), 0, s));
        HIP_OK(c, gsx::launch_transpose(2, 0, nullptr, nullptr, nullptr, nullptr, dp<int>(C.rrp), nullptr,
                                        dp<int>(C.rci), dp<double>(C.rv), nc, s));
        // B = A P with the narrow rows first, the wide ones if those overflow
        KScope kg(c, "mg_galerkin");
        int width_class = 2, errh[2] = {0, 0};
        DevPtr bcnt, bci, bv;
        for (;;) {
            const int WB = gsx::ap_width(width_class);
            bcnt = dalloc(c, sizeof(int) * std::max(1, nf));
            bci = dalloc(c, sizeof(int) * (size_t)std::max(1, nf) * WB);
            bv = dalloc(c, sizeof(double) * (size_t)std::max(1, nf) * WB);
            if (!bcnt || !bci || !bv) return fail(c, GSX_ERR_OOM, "device allocation (Galerkin product)");
            HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int) * 2, s));
            HIP_OK(c, gsx::launch_ap(width_class, nf, dp<int>(F.rowptr), dp<int>(F.col), dp<double>(F.val),
                                     dp<int>(C.prp), dp<int>(C.pci), dp<double>(C.pv), dp<int>(bcnt), dp<int>(bci),
                                     dp<double>(bv), dp<int>(errf), s));
            DevPtr ccount = dalloc(c, sizeof(int) * (nc + 1));
            if (!ccount) return fail(c, GSX_ERR_OOM, "device allocation (Galerkin pattern)");
            HIP_OK(c, gsx::launch_rap_pattern(width_class, nc, dp<int>(C.rrp), dp<int>(C.rci), dp<int>(bcnt),
                                              dp<int>(bci), dp<int>(ccount), nullptr, nullptr, dp<int>(errf), s));
            HIP_OK(c, hipMemcpyAsync(errh, errf.get(), sizeof(int) * 2, hipMemcpyDeviceToHost, s));
            HIP_OK(c, hipStreamSynchronize(s));
            if (errh[0] == 0) {
                int64_t cn = 0;
                GSX_TRY(scan_counts(c, dp<int>(ccount), nc, C.rowptr, cn));
                if (cn > INT32_MAX) return fail(c, GSX_ERR_ARG, "coarse operator too large for int32 indices");
                C.nnz = cn;
                C.col = dalloc(c, sizeof(int) * std::max<int64_t>(1, cn));
                C.val = dalloc(c, sizeof(double) * std::max<int64_t>(1, cn));
                if (!C.col || !C.val) return fail(c, GSX_ERR_OOM, "device allocation (coarse operator)");
                HIP_OK(c, gsx::launch_rap_pattern(width_class, nc, dp<int>(C.rrp), dp<int>(C.rci), dp<int>(bcnt),
                                                  dp<int>(bci), nullptr, dp<int>(C.rowptr), dp<int>(C.col),
                                                  dp<int>(errf), s));
                HIP_OK(c, gsx::launch_rap(width_class, nc, dp<int>(C.rrp), dp<int>(C.rci), dp<double>(C.rv),
                                          dp<int>(bcnt), dp<int>(bci), dp<double>(bv), dp<int>(C.rowptr),
                                          dp<int>(C.col), dp<double>(C.val), dp<int>(errf) + 1, s));
                HIP_OK(c, hipMemcpyAsync(errh, errf.get(), sizeof(int) * 2, hipMemcpyDeviceToHost, s));
                HIP_OK(c, hipStreamSynchronize(s));
                if (errh[0] == 0 && errh[1] == 0) break;
            }
            if (width_class == 3) {
                char buf[160];
                std::snprintf(buf, sizeof(buf), "multigrid level %d: Galerkin product rows too wide (codes %d, %d)", l,
                              errh[0], errh[1]);
                return fail(c, GSX_ERR_ARG, buf);
            }
            width_class = 3;
        }
        C.lpr = gsx::spmv_lpr(C.n ? (double)C.nnz / C.n : 1.0);
    }
    GSX_TRY(mg_finish(mg.get()));
    *out = mg.release();
    return GSX_OK;
}

// ---- p-multigrid (examples/pMultiGrid_example.cpp:296-560) ----------------

// Lumped mass L(i) = int phi_i dx over the free DOFs of U, with U's own
// quadrature and the geometry of G (assembleLumpedMass,
// examples/pMultiGrid_example.cpp:803-826).
static gsx_status pmg_lumped(gsx_ctx *c, gsx_problem *U, gsx_problem *G, DevPtr &L)
{
    hipStream_t s = c->stream;
    DevPtr Fb = dalloc(c, sizeof(double) * std::max<int64_t>(1, U->fdoubles));
    L = dalloc(c, sizeof(double) * std::max(1, U->nfree));
    if (!Fb || !L) return fail(c, GSX_ERR_OOM, "device allocation (lumped mass)");
    for (int ip = 0; ip < U->npatch; ++ip)
        HIP_OK(c, gsx::launch_geo_elem(U->d, dp<PatchDev>(U->d_patches), dp<PatchDev>(G->d_patches), ip, 0,
                                       U->hp[ip].nelem, c->gt, nullptr, dp<double>(Fb), nullptr, 0, 0, 1, s));
    HIP_OK(c, gsx::launch_vec_gather(U->d, dp<PatchDev>(U->d_patches), dp<int>(U->d_occ_ptr), dp<int>(U->d_occ_patch),
                                     dp<int64_t>(U->d_occ_local), U->nfree, dp<double>(Fb), dp<double>(L), s));
    return GSX_OK;
}

// Mixed mass M(i, j) = int phi_i psi_j dx, rows the free DOFs of U (the finer
// level), columns those of V, integrated on U's elements with U's quadrature
// over the geometry of G (assembleMixedMass, pMultiGrid_example.cpp:855-880,
// whose matrix().transpose() is this U x V matrix), into C.prp/pci/pv.
static gsx_status pmg_mixed(gsx_ctx *c, gsx_problem *U, gsx_problem *V, gsx_problem *G, MGLevel &C)
{
    hipStream_t s = c->stream;
    const int d = U->d, nrows = U->nfree;
    std::vector<int64_t> moff((size_t)U->npatch);
    int64_t total = 0;
    for (int k = 0; k < U->npatch; ++k) {
        const HostPatch &hu = U->hp[(size_t)k], &hv = V->hp[(size_t)k];
        int64_t nbu = 1, nbv = 1;
        for (int j = 0; j < d; ++j) {
            nbu *= hu.p[j] + 1;
            nbv *= hv.p[j] + 1;
            // V's breakpoints must be breakpoints of U
            const std::vector<double> &ku = U->knots_h[(size_t)3 * k + j], &kv = V->knots_h[(size_t)3 * k + j];
            for (double v : kv)
                if (!std::binary_search(ku.begin(), ku.end(), v))
                    return fail(c, GSX_ERR_ARG, "p-multigrid: a coarser level's mesh must be nested in the finer one's");
        }
        moff[(size_t)k] = total;
        total += hu.nelem * nbu * nbv;
    }
    DevPtr Mb = dalloc(c, sizeof(double) * std::max<int64_t>(1, total));
    DevPtr d_moff = dalloc(c, sizeof(int64_t) * moff.size());
    if (!Mb || !d_moff) return fail(c, GSX_ERR_OOM, "device allocation (mixed mass)");
    HIP_OK(c, hipMemcpyAsync(d_moff.get(), moff.data(), sizeof(int64_t) * moff.size(), hipMemcpyHostToDevice, s));
    {
        KScope k(c, "pmg_mixed_elem");
        for (int ip = 0; ip < U->npatch; ++ip)
            HIP_OK(c, gsx::launch_geo_mixed(d, dp<PatchDev>(U->d_patches), dp<PatchDev>(V->d_patches),
                                            dp<PatchDev>(G->d_patches), ip, U->hp[(size_t)ip].nelem, c->gt,
                                            dp<double>(Mb), moff[(size_t)ip], s));
    }
    // pattern: candidates per row, sorted and made unique, then compacted
    DevPtr counts = dalloc(c, sizeof(int) * (nrows + 1));
    if (!counts) return fail(c, GSX_ERR_OOM, "device allocation (mixed mass pattern)");
    const PatchDev *ups = dp<PatchDev>(U->d_patches), *vps = dp<PatchDev>(V->d_patches);
    HIP_OK(c, gsx::launch_mix_pattern(0, d, ups, vps, dp<int>(U->d_occ_ptr), dp<int>(U->d_occ_patch),
                                      dp<int64_t>(U->d_occ_local), nrows, V->nfree, nullptr, nullptr, dp<int>(counts),
                                      nullptr, nullptr, s));
    DevPtr cptr;
    int64_t ncand = 0;
    GSX_TRY(scan_counts(c, dp<int>(counts), nrows, cptr, ncand));
    if (ncand > INT32_MAX) return fail(c, GSX_ERR_ARG, "mixed mass too large for int32 indices");
    DevPtr cand = dalloc(c, sizeof(int) * std::max<int64_t>(1, ncand));
    if (!cand) return fail(c, GSX_ERR_OOM, "device allocation (mixed mass pattern)");
    HIP_OK(c, gsx::launch_mix_pattern(1, d, ups, vps, dp<int>(U->d_occ_ptr), dp<int>(U->d_occ_patch),
                                      dp<int64_t>(U->d_occ_local), nrows, V->nfree, dp<int>(cptr), dp<int>(cand),
                                      dp<int>(counts), nullptr, nullptr, s));
    int64_t nnz = 0;
    GSX_TRY(scan_counts(c, dp<int>(counts), nrows, C.prp, nnz));
    C.pci = dalloc(c, sizeof(int) * std::max<int64_t>(1, nnz));
    C.pv = dalloc(c, sizeof(double) * std::max<int64_t>(1, nnz));
    if (!C.pci || !C.pv) return fail(c, GSX_ERR_OOM, "device allocation (mixed mass)");
    HIP_OK(c, gsx::launch_mix_pattern(2, d, ups, vps, nullptr, nullptr, nullptr, nrows, V->nfree, dp<int>(cptr),
                                      dp<int>(cand), nullptr, dp<int>(C.prp), dp<int>(C.pci), s));
    {
        // one wavefront per row pays where an element has at least 27 coarse
        // functions (the lanes take them); it holds rows of up to 1024
        int nbv = 1;
        for (int j = 0; j < d; ++j) nbv *= V->hp[0].p[j] + 1;
        int maxrow = INT32_MAX;
        if (nbv >= 27) {
            std::vector<int> rp((size_t)nrows + 1);
            HIP_OK(c, hipMemcpyAsync(rp.data(), C.prp.get(), sizeof(int) * (nrows + 1), hipMemcpyDeviceToHost, s));
            HIP_OK(c, hipStreamSynchronize(s));
            maxrow = 0;
            for (int i = 0; i < nrows; ++i) maxrow = std::max(maxrow, rp[(size_t)i + 1] - rp[(size_t)i]);
        }
        KScope k(c, "pmg_mixed_gather");
        HIP_OK(c, gsx::launch_mix_gather(d, ups, vps, dp<int>(U->d_occ_ptr), dp<int>(U->d_occ_patch),
                                         dp<int64_t>(U->d_occ_local), nrows, V->nfree, dp<int>(C.prp),
                                         dp<int>(C.pci), dp<double>(Mb), dp<int64_t>(d_moff), dp<double>(C.pv),
                                         maxrow, s));
    }
    C.nfine = nrows;
    C.pnnz = nnz;
    return GSX_OK;
}

// Open knot vector of degree q over the breakpoints of U (simple interior
// knots: the maximal smoothness of gsKnotVector::degreeIncrease/
// degreeReduce on such a vector).
static std::vector<double> pmg_knots(const std::vector<double> &U, int q)
{
    std::vector<double> bp;
    for (double u : U)
        if (bp.empty() || u != bp.back()) bp.push_back(u);
    std::vector<double> K((size_t)q, bp.front());
    K.insert(K.end(), bp.begin(), bp.end());
    K.insert(K.end(), (size_t)q, bp.back());
    return K;
}

extern "C" gsx_status gsx_pmg_create(gsx_problem *fine, gsx_matrix *A, int32_t nlevels, const char *coarsening,
                                     int32_t projection, int32_t coarse_operator, int32_t npre, int32_t npost,
                                     double damping, double cgc_damping, int32_t precision, gsx_mg **out)
{
    if (!fine || !A || !out || !coarsening || nlevels < 1 || npre < 0 || npost < 0) return GSX_ERR_ARG;
    gsx_ctx *c = fine->ctx;
    if (fine->geo) return fail(c, GSX_ERR_ARG, "the finest problem must carry its own geometry");
    if (A->n != fine->nfree) return fail(c, GSX_ERR_ARG, "matrix does not belong to the problem");
    if (precision != GSX_FP64 && precision != GSX_MIXED) return fail(c, GSX_ERR_ARG, "unknown precision");
    if (projection != 1 && projection != 2) return fail(c, GSX_ERR_ARG, "projection must be 1 or 2");
    if (coarse_operator != 1 && coarse_operator != 2) return fail(c, GSX_ERR_ARG, "coarse operator must be 1 or 2");
    if ((int)std::strlen(coarsening) != nlevels - 1)
        return fail(c, GSX_ERR_ARG, "coarsening needs one of h, p, z per level transition (nlevels - 1)");
    int npz = 0;
    for (int i = 0; i < nlevels - 1; ++i) {
        const char t = coarsening[i];
        if (t != 'h' && t != 'p' && t != 'z') return fail(c, GSX_ERR_ARG, "coarsening: unknown type (h, p or z)");
        npz += t != 'h';
    }
    if (projection == 1 && npz > 1)
        return fail(c, GSX_ERR_ARG, "projection 1 allows only one p or z transition");
    hipStream_t s = c->stream;
    KScope ks(c, "pmg_setup");
    const int npatch = fine->npatch, d = fine->d;
    // level bases, finest first: knots per (patch, direction) and degree per patch
    std::vector<std::vector<std::vector<double>>> knots((size_t)nlevels);
    std::vector<std::vector<int>> deg((size_t)nlevels);
    knots[(size_t)nlevels - 1] = fine->knots_h;
    for (int k = 0; k < npatch; ++k) deg[(size_t)nlevels - 1].push_back(fine->hp[(size_t)k].p[0]);
    for (int l = nlevels - 2; l >= 0; --l) {
        const char t = coarsening[l];
        knots[(size_t)l].resize((size_t)npatch * 3);
        deg[(size_t)l].resize((size_t)npatch);
        for (int k = 0; k < npatch; ++k) {
            const int pf = deg[(size_t)l + 1][(size_t)k];
            int pc = pf;
            if (t != 'h') {
                pc = projection == 1 ? 1 : pf - 1;
                if (pc < 1 || pc >= pf) return fail(c, GSX_ERR_ARG, "p-multigrid: degree cannot be reduced further");
            }
            deg[(size_t)l][(size_t)k] = pc;
            for (int j = 0; j < d; ++j) {
                std::vector<double> U = knots[(size_t)l + 1][(size_t)3 * k + j];
                if (t != 'p') {
                    const int nel = (int)U.size() - 2 * pf - 1;
                    if (nel % 2 != 0 || nel < 2)
                        return fail(c, GSX_ERR_ARG,
                                    "multigrid needs an even number of elements per direction on every level");
                    std::vector<double> V(U.begin(), U.begin() + pf + 1);
                    for (int e = 2; e < nel; e += 2) V.push_back(U[(size_t)(pf + e)]);
                    V.insert(V.end(), U.end() - pf - 1, U.end());
                    U.swap(V);
                }
                knots[(size_t)l][(size_t)3 * k + j] = pmg_knots(U, pc);
            }
        }
    }
    auto mg = std::make_unique<gsx_mg>();
    mg->ctx = c;
    mg->npre = npre;
    mg->npost = npost;
    mg->ncycles = 1;
    mg->tau = damping;
    mg->cgc = cgc_damping;
    mg->precision = precision;
    mg->L.resize((size_t)nlevels);
    mg->fine_box_ok = scms_box_of(fine, mg->fine_box);
    mg->fine_layout = layout_of(fine);
    mg->fine_layout_ok = true;
    {
        MGLevel &f = mg->L[(size_t)nlevels - 1];
        f.n = A->n;
        f.nnz = A->nnz;
        f.lpr = A->lpr;
        f.rowptr = A->rowptr;
        f.col = A->col;
        f.val = A->val;
    }
    // level problems over the fine geometry
    for (int l = nlevels - 2; l >= 0; --l) {
        std::vector<gsx_patch_desc> desc((size_t)npatch);
        for (int k = 0; k < npatch; ++k) {
            gsx_patch_desc &D = desc[(size_t)k];
            std::memset(&D, 0, sizeof(D));
            D.d = d;
            for (int j = 0; j < d; ++j) {
                D.p[j] = deg[(size_t)l][(size_t)k];
                D.nknots[j] = (int)knots[(size_t)l][(size_t)3 * k + j].size();
                D.knots_host[j] = knots[(size_t)l][(size_t)3 * k + j].data();
            }
            D.fkind = 1;
        }
        gsx_problem *cp = nullptr;
        GSX_TRY(gsx_problem_create_geo(fine, npatch, desc.data(), (int)fine->dir_sides_h.size() / 2,
                                       fine->dir_sides_h.empty() ? nullptr : fine->dir_sides_h.data(),
                                       (int)fine->ifaces_h.size() / 8,
                                       fine->ifaces_h.empty() ? nullptr : fine->ifaces_h.data(), &cp));
        mg->L[(size_t)l].prob = cp;
        mg->L[(size_t)l].n = cp->nfree;
    }
    // operators and transfers, finest first
    TransferCache tcache;
    for (int l = nlevels - 2; l >= 0; --l) {
        MGLevel &C = mg->L[(size_t)l], &F = mg->L[(size_t)l + 1];
        gsx_problem *fp = (l + 1 == nlevels - 1) ? fine : F.prob;
        gsx_problem *cp = C.prob;
        const char t = coarsening[l];
        const bool galerkin = t == 'h' && coarse_operator == 2;
        if (galerkin) {
            if (!cp->have_pattern) GSX_TRY(pattern_build(cp));
            C.rowptr = cp->d_rowptr;
            C.col = cp->d_col;
            C.nnz = cp->nnz;
            C.val = dalloc(c, sizeof(double) * std::max<int64_t>(1, C.nnz));
            if (!C.val) return fail(c, GSX_ERR_OOM, "device allocation (Galerkin product)");
        } else {
            // rediscretisation on the level's basis over the fine geometry
            gsx_matrix *Al = nullptr;
            GSX_TRY(gsx_assemble(cp, GSX_PATTERN_FULL | GSX_ASM_CSR_ONLY, GSX_FP64, &Al, nullptr));
            C.rowptr = Al->rowptr;
            C.col = Al->col;
            C.val = Al->val;
            C.nnz = Al->nnz;
            gsx_matrix_destroy(Al);
        }
        C.lpr = gsx::spmv_lpr(C.n ? (double)C.nnz / C.n : 1.0);
        KScope kt(c, "pmg_transfer_build");
        if (t == 'h') {
            // canonical embedding and its transpose
            GSX_TRY(mg_knot_transfer(c, fp, cp, F.n, C, tcache, nullptr, nullptr));
            GSX_TRY(mg_transpose_transfer(c, C, F.n));
        } else {
            // L2 projections with lumped mass: P = L_f^-1 M, R = L_c^-1 M^T
            DevPtr Lf, Lc;
            GSX_TRY(pmg_mixed(c, fp, cp, fine, C));
            GSX_TRY(mg_transpose_transfer(c, C, F.n));
            GSX_TRY(pmg_lumped(c, fp, fine, Lf));
            GSX_TRY(pmg_lumped(c, cp, fine, Lc));
            DevPtr errf = dalloc(c, sizeof(int));
            if (!errf) return fail(c, GSX_ERR_OOM, "device allocation");
            HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int), s));
            HIP_OK(c, gsx::launch_scale_rows(F.n, dp<int>(C.prp), dp<double>(Lf), dp<double>(C.pv), dp<int>(errf), s));
            HIP_OK(c, gsx::launch_scale_rows(C.n, dp<int>(C.rrp), dp<double>(Lc), dp<double>(C.rv), dp<int>(errf), s));
            int errh = 0;
            HIP_OK(c, hipMemcpyAsync(&errh, errf.get(), sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_OK(c, hipStreamSynchronize(s));
            if (errh) return fail(c, GSX_ERR_ARG, "p-multigrid: a lumped mass entry is zero");
        }
        if (galerkin) {
            DevPtr errf = dalloc(c, sizeof(int) * 2);
            if (!errf) return fail(c, GSX_ERR_OOM, "device allocation");
            HIP_OK(c, hipMemsetAsync(errf.get(), 0, sizeof(int) * 2, s));
            GSX_TRY(mg_galerkin_sparse(c, C, F, d, dp<int>(errf)));
            int errh[2] = {0, 0};
            HIP_OK(c, hipMemcpyAsync(errh, errf.get(), sizeof(int) * 2, hipMemcpyDeviceToHost, s));
            HIP_OK(c, hipStreamSynchronize(s));
            if (errh[0] || errh[1]) return fail(c, GSX_ERR_ARG, "p-multigrid: Galerkin product buffer overflow");
        }
    }
    GSX_TRY(mg_finish(mg.get()));
    *out = mg.release();
    return GSX_OK;
}

extern "C" gsx_status gsx_mg_destroy(gsx_mg *mg)
{
    if (!mg) return GSX_ERR_ARG;
    (void)hipStreamSynchronize(mg->ctx->stream);
    delete mg;
    return GSX_OK;
}

extern "C" gsx_status gsx_mg_level_info(const gsx_mg *mg, int32_t lvl, int64_t *n, int64_t *nnz, int64_t *pnnz)
{
    if (!mg || lvl < 0 || lvl >= (int)mg->L.size()) return GSX_ERR_ARG;
    const MGLevel &l = mg->L[lvl];
    if (n) *n = l.n;
    if (nnz) *nnz = l.nnz;
    if (pnnz) *pnnz = l.pnnz;
    return GSX_OK;
}

// level operator (CSR) and, for lvl < nlevels-1, the transfer to the next
// finer level (CSR, n_fine rows)
extern "C" gsx_status gsx_mg_level_download(const gsx_mg *mg, int32_t lvl, int32_t *rp, int32_t *ci, double *v,
                                            int32_t *prp, int32_t *pci, double *pv)
{
    if (!mg || lvl < 0 || lvl >= (int)mg->L.size()) return GSX_ERR_ARG;
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    const MGLevel &l = mg->L[lvl];
    if (rp) HIP_OK(c, hipMemcpyAsync(rp, l.rowptr.get(), sizeof(int) * (l.n + 1), hipMemcpyDeviceToHost, s));
    if (ci) HIP_OK(c, hipMemcpyAsync(ci, l.col.get(), sizeof(int) * l.nnz, hipMemcpyDeviceToHost, s));
    if (v) HIP_OK(c, hipMemcpyAsync(v, l.val.get(), sizeof(double) * l.nnz, hipMemcpyDeviceToHost, s));
    if (lvl < (int)mg->L.size() - 1) {
        if (prp) HIP_OK(c, hipMemcpyAsync(prp, l.prp.get(), sizeof(int) * (l.nfine + 1), hipMemcpyDeviceToHost, s));
        if (pci) HIP_OK(c, hipMemcpyAsync(pci, l.pci.get(), sizeof(int) * l.pnnz, hipMemcpyDeviceToHost, s));
        if (pv) HIP_OK(c, hipMemcpyAsync(pv, l.pv.get(), sizeof(double) * l.pnnz, hipMemcpyDeviceToHost, s));
    }
    HIP_OK(c, hipStreamSynchronize(s));
    return GSX_OK;
}

extern "C" gsx_status gsx_mg_restriction_download(const gsx_mg *mg, int32_t lvl, int32_t *rp, int32_t *ci, double *v)
{
    if (!mg || lvl < 0 || lvl >= (int)mg->L.size() - 1) return GSX_ERR_ARG;
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    const MGLevel &l = mg->L[lvl];
    if (rp) HIP_OK(c, hipMemcpyAsync(rp, l.rrp.get(), sizeof(int) * (l.n + 1), hipMemcpyDeviceToHost, s));
    if (ci) HIP_OK(c, hipMemcpyAsync(ci, l.rci.get(), sizeof(int) * l.pnnz, hipMemcpyDeviceToHost, s));
    if (v) HIP_OK(c, hipMemcpyAsync(v, l.rv.get(), sizeof(double) * l.pnnz, hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    return GSX_OK;
}

// One multigrid step at level l on (b, x) with x = 0 on entry (the
// preconditioner's apply).  FP64 version.
// Polynomial smoother coefficients of sweep k (Chebyshev recurrence on
// [lmax / ratio, lmax]; Richardson: constant tau, no diagonal).
struct SmoothCoef {
    double c1, c2;
};
static SmoothCoef smooth_coef(const gsx_mg *mg, const MGLevel &V, int k, double &rho)
{
    (void)mg;
    if (V.smoother == GSX_SMOOTH_RICHARDSON) return {0.0, V.tau};
    const double l1 = V.lmax, l0 = l1 / V.cheb_ratio;
    const double theta = 0.5 * (l1 + l0), delta = 0.5 * (l1 - l0), sigma = theta / delta;
    if (k == 0) {
        rho = 1.0 / sigma;
        return {0.0, 1.0 / theta};
    }
    const double rn = 1.0 / (2.0 * sigma - rho);
    const SmoothCoef c{rn * rho, 2.0 * rn / delta};
    rho = rn;
    return c;
}

// m smoothing sweeps on level V with FP64 vectors (matrix values FP64, or
// FP32 when v32: the finest level of a mixed cycle).  Jacobi ping-pongs x
// and t (x may come back as t); the polynomial smoothers update x in place.
static gsx_status smooth64(gsx_mg *mg, MGLevel &V, double *&x, double *&t, const double *b, bool &zero, int m,
                           bool v32, bool post)
{
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    const int *rp = dp<int>(V.rowptr), *ci = dp<int>(V.col);
    const double *va = dp<double>(V.val), *dg = dp<double>(V.diag);
    const float *va32 = v32 ? dp<float>(V.val32) : nullptr;
    auto residual = [&](const double *xx, double *rr) {
        return v32 ? gsx::launch_residual_v32(V.lpr, V.n, rp, ci, va32, xx, b, rr, s)
                   : gsx::launch_residual_f64(V.lpr, V.n, rp, ci, va, xx, b, rr, s);
    };
    if (V.smoother == GSX_SMOOTH_GAUSS_SEIDEL) {
        if (m > 0 && zero) HIP_OK(c, gsx::launch_fill_f64(x, V.n, 0.0, s));
        const int *rows = dp<int>(post ? V.gs_rev : V.gs_fwd);
        const std::vector<int> &off = post ? V.gs_rev_off : V.gs_fwd_off;
        for (int k = 0; k < m; ++k)
            for (size_t l = 0; l + 1 < off.size(); ++l) {
                if (v32)
                    HIP_OK(c, gsx::launch_gs_rows_v32(off[l + 1] - off[l], rows + off[l], rp, ci, va32, b, x, s));
                else
                    HIP_OK(c, gsx::launch_gs_rows_f64(off[l + 1] - off[l], rows + off[l], rp, ci, va, b, x, s));
            }
        if (m > 0) zero = false;
        return GSX_OK;
    }
    if (V.smoother == GSX_SMOOTH_SCMS && V.split) {
        // x += sum_c T_c op_c T_c^T (b - A x): the components partition the
        // DOFs, so their updates do not overlap
        if (m > 0 && zero) HIP_OK(c, gsx::launch_fill_f64(x, V.n, 0.0, s));
        double *r = dp<double>(V.r);
        ScmsSplit &S = *V.split;
        double *rc = dp<double>(S.rc), *zc = dp<double>(S.zc);
        for (int k = 0; k < m; ++k) {
            HIP_OK(c, residual(x, r));
            for (const ScmsComp &q : S.comps) {
                HIP_OK(c, gsx::launch_gather_idx(q.n, dp<int>(q.idx), r, rc, s));
                if (q.F) {
                    const gsx_fd *F = q.F.get();
                    HIP_OK(c, fd_apply<double>(F, rc, 0, F->m[0], (int64_t)F->m[0] * F->m[1], zc, 0, F->m[0],
                                               (int64_t)F->m[0] * F->m[1], nullptr, s));
                } else {
                    HIP_OK(c, gsx::launch_dense_gemv_f64(q.n, q.n, dp<double>(q.Ainv), rc, zc, s));
                }
                HIP_OK(c, gsx::launch_scatter_add_idx(q.n, dp<int>(q.idx), zc, q.tau, x, s));
            }
        }
        if (m > 0) zero = false;
        return GSX_OK;
    }
    if (V.smoother == GSX_SMOOTH_SCMS) {
        // x += tau S (b - A x)
        if (m > 0 && zero) HIP_OK(c, gsx::launch_fill_f64(x, V.n, 0.0, s));
        double *r = dp<double>(V.r), *d = dp<double>(V.d);
        const gsx_fd *F = V.scms.get();
        for (int k = 0; k < m; ++k) {
            HIP_OK(c, residual(x, r));
            HIP_OK(c, fd_apply<double>(F, r, 0, F->m[0], (int64_t)F->m[0] * F->m[1], d, 0, F->m[0],
                                       (int64_t)F->m[0] * F->m[1], nullptr, s));
            HIP_OK(c, gsx::launch_smooth_update_f64(V.n, x, d, d, nullptr, 0.0, V.scms_tau, 1, s));
        }
        if (m > 0) zero = false;
        return GSX_OK;
    }
    if (V.smoother == GSX_SMOOTH_JACOBI) {
        for (int i = 0; i < m; ++i) {
            if (v32)
                HIP_OK(c, gsx::launch_jacobi_step_v32(V.lpr, V.n, rp, ci, va32, dg, b, x, t, V.tau, zero ? 1 : 0, s));
            else
                HIP_OK(c, gsx::launch_jacobi_step_f64(V.lpr, V.n, rp, ci, va, dg, b, x, t, V.tau, zero ? 1 : 0, s));
            std::swap(x, t);
            zero = false;
        }
        return GSX_OK;
    }
    if (m > 0 && zero) HIP_OK(c, gsx::launch_fill_f64(x, V.n, 0.0, s));
    double *r = dp<double>(V.r), *d = dp<double>(V.d);
    double rho = 0.0;
    for (int k = 0; k < m; ++k) {
        HIP_OK(c, residual(x, r));
        const SmoothCoef cf = smooth_coef(mg, V, k, rho);
        HIP_OK(c, gsx::launch_smooth_update_f64(V.n, x, d, r, V.smoother == GSX_SMOOTH_CHEBYSHEV ? dg : nullptr,
                                                cf.c1, cf.c2, k == 0 ? 1 : 0, s));
        zero = false;
    }
    return GSX_OK;
}

static gsx_status smooth32(gsx_mg *mg, MGLevel &V, float *&x, float *&t, const float *b, bool &zero, int m,
                           bool post)
{
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    const int *rp = dp<int>(V.rowptr), *ci = dp<int>(V.col);
    const float *va = dp<float>(V.val32), *dg = dp<float>(V.diag32);
    if (V.smoother == GSX_SMOOTH_GAUSS_SEIDEL) {
        if (m > 0 && zero) HIP_OK(c, gsx::launch_fill_f32(x, V.n, 0.0f, s));
        const int *rows = dp<int>(post ? V.gs_rev : V.gs_fwd);
        const std::vector<int> &off = post ? V.gs_rev_off : V.gs_fwd_off;
        for (int k = 0; k < m; ++k)
            for (size_t l = 0; l + 1 < off.size(); ++l)
                HIP_OK(c, gsx::launch_gs_rows_f32(off[l + 1] - off[l], rows + off[l], rp, ci, va, b, x, s));
        if (m > 0) zero = false;
        return GSX_OK;
    }
    if (V.smoother == GSX_SMOOTH_SCMS) {
        if (m > 0 && zero) HIP_OK(c, gsx::launch_fill_f32(x, V.n, 0.0f, s));
        float *r = dp<float>(V.r32), *d = dp<float>(V.d32);
        const gsx_fd *F = V.scms.get();
        for (int k = 0; k < m; ++k) {
            HIP_OK(c, gsx::launch_residual_f32(V.lpr, V.n, rp, ci, va, x, b, r, s));
            HIP_OK(c, fd_apply<float>(F, r, 0, F->m[0], (int64_t)F->m[0] * F->m[1], d, 0, F->m[0],
                                      (int64_t)F->m[0] * F->m[1], nullptr, s));
            HIP_OK(c, gsx::launch_smooth_update_f32(V.n, x, d, d, nullptr, 0.0f, (float)V.scms_tau, 1, s));
        }
        if (m > 0) zero = false;
        return GSX_OK;
    }
    if (V.smoother == GSX_SMOOTH_JACOBI) {
        for (int i = 0; i < m; ++i) {
            HIP_OK(c, gsx::launch_jacobi_step_f32(V.lpr, V.n, rp, ci, va, dg, b, x, t, (float)V.tau, zero ? 1 : 0, s));
            std::swap(x, t);
            zero = false;
        }
        return GSX_OK;
    }
    if (m > 0 && zero) HIP_OK(c, gsx::launch_fill_f32(x, V.n, 0.0f, s));
    float *r = dp<float>(V.r32), *d = dp<float>(V.d32);
    double rho = 0.0;
    for (int k = 0; k < m; ++k) {
        HIP_OK(c, gsx::launch_residual_f32(V.lpr, V.n, rp, ci, va, x, b, r, s));
        const SmoothCoef cf = smooth_coef(mg, V, k, rho);
        HIP_OK(c, gsx::launch_smooth_update_f32(V.n, x, d, r, V.smoother == GSX_SMOOTH_CHEBYSHEV ? dg : nullptr,
                                                (float)cf.c1, (float)cf.c2, k == 0 ? 1 : 0, s));
        zero = false;
    }
    return GSX_OK;
}

static gsx_status mg_cycle32(gsx_mg *mg, int l);

// One V-cycle step at level l with FP64 vectors.  mixed: level l is the
// finest level of a mixed-precision cycle: its matrix values are read in
// FP32 and the coarser levels run mg_cycle32.  The fine vectors stay FP64:
// an FP32 copy of the smooth fine-level iterate carries a rounding error
// whose residual is of order eps cond(A) ||b||, which costs outer
// iterations; FP32 matrix values on their own do not.
static gsx_status mg_cycle64(gsx_mg *mg, int l, bool mixed = false)
{
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    MGLevel &V = mg->L[l];
    double *x = dp<double>(V.x), *t = dp<double>(V.t), *b = dp<double>(V.b), *r = dp<double>(V.r);
    if (l == 0) {
        KScope k(c, "mg_coarse");
        if (mg->band_b > 0)
            HIP_OK(c, gsx::launch_band_solve(V.n, mg->band_b, dp<int>(mg->band_pos), dp<double>(mg->band_L), b, x,
                                             dp<double>(mg->band_y), s));
        else
            HIP_OK(c, gsx::launch_dense_gemv_f64(V.n, mg->npad, dp<double>(mg->Dinv), b, x, s));
        return GSX_OK;
    }
    const int *rp = dp<int>(V.rowptr), *ci = dp<int>(V.col);
    const double *va = dp<double>(V.val);
    bool zero = true;
    {
        KScope k(c, "mg_smooth");
        GSX_TRY(smooth64(mg, V, x, t, b, zero, mg->npre, mixed, false));
    }
    MGLevel &C = mg->L[l - 1];
    {
        KScope k(c, "mg_transfer");
        if (zero) HIP_OK(c, hipMemcpyAsync(r, b, sizeof(double) * V.n, hipMemcpyDeviceToDevice, s));
        else if (mixed) HIP_OK(c, gsx::launch_residual_v32(V.lpr, V.n, rp, ci, dp<float>(V.val32), x, b, r, s));
        else HIP_OK(c, gsx::launch_residual_f64(V.lpr, V.n, rp, ci, va, x, b, r, s));
        if (mixed)
            HIP_OK(c, gsx::launch_restrict_v32(C.n, dp<int>(C.rrp), dp<int>(C.rci), dp<float>(C.rv32), r,
                                               dp<float>(C.b32), s));
        else
            HIP_OK(c, gsx::launch_restrict_f64(C.n, dp<int>(C.rrp), dp<int>(C.rci), dp<double>(C.rv), r,
                                               dp<double>(C.b), s));
    }
    // coarse correction from a zero start (one V-cycle step)
    if (mixed) GSX_TRY(mg_cycle32(mg, l - 1));
    else GSX_TRY(mg_cycle64(mg, l - 1));
    {
        KScope k(c, "mg_transfer");
        if (zero) HIP_OK(c, gsx::launch_fill_f64(x, V.n, 0.0, s));
        if (mixed)
            HIP_OK(c, gsx::launch_prolong_add_v32(V.n, dp<int>(C.prp), dp<int>(C.pci), dp<float>(C.pv32),
                                                  dp<float>(C.x32), x, mg->cgc, s));
        else
            HIP_OK(c, gsx::launch_prolong_add_f64(V.n, dp<int>(C.prp), dp<int>(C.pci), dp<double>(C.pv),
                                                  dp<double>(C.x), x, mg->cgc, s));
    }
    {
        KScope k(c, "mg_smooth");
        GSX_TRY(smooth64(mg, V, x, t, b, zero, mg->npost, mixed, true));
    }
    if (x != dp<double>(V.x)) HIP_OK(c, hipMemcpyAsync(dp<double>(V.x), x, sizeof(double) * V.n, hipMemcpyDeviceToDevice, s));
    return GSX_OK;
}

// FP32 cycle below the finest level of a mixed cycle (operators, transfers
// and vectors in FP32; the coarsest solve in FP64 through conversions).
static gsx_status mg_cycle32(gsx_mg *mg, int l)
{
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    MGLevel &V = mg->L[l];
    float *x = dp<float>(V.x32), *t = dp<float>(V.t32), *b = dp<float>(V.b32), *r = dp<float>(V.r32);
    if (l == 0) {
        KScope k(c, "mg_coarse");
        HIP_OK(c, gsx::launch_conv_f32_f64(V.n, b, dp<double>(V.b), s));
        if (mg->band_b > 0)
            HIP_OK(c, gsx::launch_band_solve(V.n, mg->band_b, dp<int>(mg->band_pos), dp<double>(mg->band_L),
                                             dp<double>(V.b), dp<double>(V.x), dp<double>(mg->band_y), s));
        else
            HIP_OK(c, gsx::launch_dense_gemv_f64(V.n, mg->npad, dp<double>(mg->Dinv), dp<double>(V.b),
                                                 dp<double>(V.x), s));
        HIP_OK(c, gsx::launch_conv_f64_f32(V.n, dp<double>(V.x), x, s));
        return GSX_OK;
    }
    const int *rp = dp<int>(V.rowptr), *ci = dp<int>(V.col);
    const float *va = dp<float>(V.val32);
    bool zero = true;
    {
        KScope k(c, "mg_smooth");
        GSX_TRY(smooth32(mg, V, x, t, b, zero, mg->npre, false));
    }
    MGLevel &C = mg->L[l - 1];
    {
        KScope k(c, "mg_transfer");
        if (zero) HIP_OK(c, hipMemcpyAsync(r, b, sizeof(float) * V.n, hipMemcpyDeviceToDevice, s));
        else HIP_OK(c, gsx::launch_residual_f32(V.lpr, V.n, rp, ci, va, x, b, r, s));
        HIP_OK(c, gsx::launch_restrict_f32(C.n, dp<int>(C.rrp), dp<int>(C.rci), dp<float>(C.rv32), r, dp<float>(C.b32), s));
    }
    GSX_TRY(mg_cycle32(mg, l - 1));
    {
        KScope k(c, "mg_transfer");
        if (zero) HIP_OK(c, gsx::launch_fill_f32(x, V.n, 0.0f, s));
        HIP_OK(c, gsx::launch_prolong_add_f32(V.n, dp<int>(C.prp), dp<int>(C.pci), dp<float>(C.pv32), dp<float>(C.x32), x,
                                              (float)mg->cgc, s));
    }
    {
        KScope k(c, "mg_smooth");
        GSX_TRY(smooth32(mg, V, x, t, b, zero, mg->npost, true));
    }
    if (x != dp<float>(V.x32)) HIP_OK(c, hipMemcpyAsync(dp<float>(V.x32), x, sizeof(float) * V.n, hipMemcpyDeviceToDevice, s));
    return GSX_OK;
}

// z = MG(r) (gsPreconditionerOp::apply: x = 0, one step)
static gsx_status mg_apply_dev(gsx_mg *mg, const double *r, double *z)
{
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    MGLevel &F = mg->L.back();
    const int nl = (int)mg->L.size();
    HIP_OK(c, hipMemcpyAsync(dp<double>(F.b), r, sizeof(double) * F.n, hipMemcpyDeviceToDevice, s));
    GSX_TRY(mg_cycle64(mg, nl - 1, mg->precision == GSX_MIXED && nl > 1));
    HIP_OK(c, hipMemcpyAsync(z, dp<double>(F.x), sizeof(double) * F.n, hipMemcpyDeviceToDevice, s));
    return GSX_OK;
}

// Components of setupSubspaceCorrectedMassSmoother
// (examples/multiGrid_example.cpp:407-567) with conforming interfaces:
// gsBoxTopology::allComponents(true) (src/gsCore/gsBoxTopology.cpp:294-340)
// splits every patch into its 3^d sub-boxes (interior, sides, edges,
// corners); sub-boxes of glued patches with the same corners are one
// component, and all corners are one.  A component's DOFs are the free
// functions of its sub-boxes whose index is interior in the sub-box's open
// directions (componentBasis_withIndices with no_lower,
// src/gsCore/gsMultiBasis.hpp:262-328), the first patch's in its
// lexicographic order, then the others' not yet listed.  Patch interiors get
// the smoother of the patch basis (sigma, alpha = 0, beta = 1), faces of a
// 3D patch that of the face basis with alpha = p/h, beta = h/p (h the
// patch's smallest knot span, p its degree); all others the exact solve of
// A restricted to the component (makeSparseCholeskySolver).
struct CompDesc {
    int dim = 0, patch = 0;
    std::vector<int> idx;
    std::vector<char> seen;
    bool box = false;
    ScmsBox B;
    int boxsize = 0;
    double alpha = 0.0, beta = 1.0;
};

static std::vector<CompDesc> scms_components(const PatchLayout &L)
{
    const int d = L.d;
    int ncode = 1;
    for (int j = 0; j < d; ++j) ncode *= 3;
    std::map<std::pair<int, std::vector<int>>, size_t> where;
    std::vector<CompDesc> comps;
    for (int k = 0; k < L.npatch; ++k) {
        const HostPatch &h = L.hp[(size_t)k];
        for (int code = 0; code < ncode; ++code) {
            int loc[3] = {1, 1, 1}, dim = 0;   // per direction 0: first index, 1: interior, 2: last
            for (int j = 0, cc = code; j < d; ++j, cc /= 3) {
                loc[j] = cc % 3;
                dim += loc[j] == 1;
            }
            // corner functions in the sub-box's closure identify it across patches
            std::vector<int> key;
            for (int cn = 0; cn < (1 << d); ++cn) {
                bool in = true;
                int64_t m = 0, stride = 1;
                for (int j = 0; j < d; ++j) {
                    const int last = (cn >> j) & 1;
                    if ((loc[j] == 0 && last) || (loc[j] == 2 && !last)) in = false;
                    m += (last ? h.nb[j] - 1 : 0) * stride;
                    stride *= h.nb[j];
                }
                if (in) key.push_back(L.dofs[(size_t)(L.boff[(size_t)k] + m)]);
            }
            std::sort(key.begin(), key.end());
            if (dim == d) key.insert(key.begin(), -1 - k);   // one interior per patch
            if (dim == 0) key.clear();                       // all corners are one component
            auto ins = where.emplace(std::make_pair(dim, key), comps.size());
            if (ins.second) {
                CompDesc cd;
                cd.dim = dim;
                cd.patch = k;
                cd.seen.assign((size_t)L.nfree, 0);
                cd.box = dim == d || (dim == d - 1 && dim >= 2);
                if (cd.box) {
                    double hmin = 1e300;
                    int pmax = 0;
                    for (int j = 0; j < d; ++j) {
                        const std::vector<double> &U = L.knots[(size_t)3 * k + j];
                        for (size_t i = 0; i + 1 < U.size(); ++i)
                            if (U[i + 1] > U[i]) hmin = std::min(hmin, U[i + 1] - U[i]);
                        pmax = std::max(pmax, h.p[j]);
                    }
                    cd.B.d = dim;
                    cd.boxsize = 1;
                    for (int j = 0, o = 0; j < d; ++j) {
                        if (loc[j] != 1) continue;
                        cd.B.p[o] = h.p[j];
                        cd.B.nb[o] = h.nb[j];
                        cd.B.lo[o] = 1;
                        cd.B.hi[o] = h.nb[j] - 2;
                        cd.B.knots[o] = L.knots[(size_t)3 * k + j];
                        cd.boxsize *= std::max(0, h.nb[j] - 2);
                        ++o;
                    }
                    if (dim < d) {
                        cd.alpha = pmax / hmin;
                        cd.beta = hmin / pmax;
                    }
                }
                comps.push_back(std::move(cd));
            }
            CompDesc &cd = comps[ins.first->second];
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            for (int j = 0; j < d; ++j) {
                lo[j] = loc[j] == 0 ? 0 : loc[j] == 2 ? h.nb[j] - 1 : 1;
                hi[j] = loc[j] == 0 ? 0 : loc[j] == 2 ? h.nb[j] - 1 : h.nb[j] - 2;
            }
            for (int z = lo[2]; z <= hi[2]; ++z)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int x = lo[0]; x <= hi[0]; ++x) {
                        const int64_t m = x + (int64_t)h.nb[0] * (y + (int64_t)h.nb[1] * z);
                        const int g = L.dofs[(size_t)(L.boff[(size_t)k] + m)];
                        if (g < L.nfree && !cd.seen[(size_t)g]) {
                            cd.seen[(size_t)g] = 1;
                            cd.idx.push_back(g);
                        }
                    }
        }
    }
    // highest dimension first, as allComponents lists them
    std::stable_sort(comps.begin(), comps.end(), [](const CompDesc &a, const CompDesc &b) { return a.dim > b.dim; });
    for (CompDesc &cd : comps) std::vector<char>().swap(cd.seen);
    return comps;
}

// The components' operators on level l: boxes get their smoother damped by
// 1 / lambda, lambda from ten power iterations of S_c A_cc from
// v0_i = 1 / (1 + i mod 5) (estimateLargestEigenvalueOfPreconditionedSystem,
// src/gsSolver/gsPreconditioner.h:121-132), kept per component for the finer
// levels under the example's rule (level above nlevels - 5, more elements
// than the degree in every direction of the component's basis); the others
// the dense inverse of A_cc.
static gsx_status scms_split_build(gsx_mg *mg, int l, const PatchLayout &L, double sigma)
{
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    const int nl = (int)mg->L.size();
    MGLevel &V = mg->L[l];
    if (L.nfree != V.n) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: level size mismatch");
    std::vector<CompDesc> comps = scms_components(L);
    if (mg->scms_saved_comp.size() != comps.size()) mg->scms_saved_comp.assign(comps.size(), 0.0);
    std::vector<int> rp((size_t)V.n + 1), ci((size_t)V.nnz);
    std::vector<double> va((size_t)V.nnz);
    HIP_OK(c, hipMemcpyAsync(rp.data(), V.rowptr.get(), sizeof(int) * (V.n + 1), hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipMemcpyAsync(ci.data(), V.col.get(), sizeof(int) * V.nnz, hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipMemcpyAsync(va.data(), V.val.get(), sizeof(double) * V.nnz, hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    auto S = std::make_shared<ScmsSplit>();
    std::vector<int> pos((size_t)V.n, -1);
    int covered = 0;
    for (size_t i = 0; i < comps.size(); ++i) {
        const CompDesc &cd = comps[i];
        const int n = (int)cd.idx.size();
        if (n == 0) continue;
        covered += n;
        ScmsComp q;
        q.n = n;
        q.dim = cd.dim;
        q.idx = dalloc(c, sizeof(int) * n);
        if (!q.idx) return fail(c, GSX_ERR_OOM, "device allocation (smoother components)");
        HIP_OK(c, hipMemcpyAsync(q.idx.get(), cd.idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        if (cd.box) {
            if (n != cd.boxsize)
                return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: a patch or face component is not a "
                                            "box of free functions (conforming interfaces only)");
            std::unique_ptr<gsx_fd> F;
            GSX_TRY(scms_build(c, cd.B, sigma, cd.alpha, cd.beta, F));
            if (F->size != n) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: component size mismatch");
            q.F = std::shared_ptr<gsx_fd>(F.release());
            double &saved = mg->scms_saved_comp[i];
            if (saved == 0.0) {
                std::vector<double> xh((size_t)n);
                for (int j = 0; j < n; ++j) xh[(size_t)j] = 1.0 / (1.0 + (double)(j % 5));
                DevPtr cv = dalloc(c, sizeof(double) * n), cw = dalloc(c, sizeof(double) * n);
                DevPtr gv = dalloc(c, sizeof(double) * V.n), gt = dalloc(c, sizeof(double) * V.n);
                if (!cv || !cw || !gv || !gt) return fail(c, GSX_ERR_OOM, "device allocation (power iteration)");
                const gsx_fd *Fp = q.F.get();
                for (int it = 0; it < 10; ++it) {
                    double nrm = 0.0;
                    for (double v : xh) nrm += v * v;
                    nrm = std::sqrt(nrm);
                    for (double &v : xh) v /= nrm;
                    HIP_OK(c, hipMemcpyAsync(cv.get(), xh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
                    HIP_OK(c, gsx::launch_fill_f64(dp<double>(gv), V.n, 0.0, s));
                    HIP_OK(c, gsx::launch_scatter_add_idx(n, dp<int>(q.idx), dp<double>(cv), 1.0, dp<double>(gv), s));
                    HIP_OK(c, gsx::launch_spmv_f64(V.lpr, V.n, dp<int>(V.rowptr), dp<int>(V.col), dp<double>(V.val),
                                                   dp<double>(gv), dp<double>(gt), nullptr, nullptr, 0, nullptr, s));
                    HIP_OK(c, gsx::launch_gather_idx(n, dp<int>(q.idx), dp<double>(gt), dp<double>(cw), s));
                    HIP_OK(c, fd_apply<double>(Fp, dp<double>(cw), 0, Fp->m[0], (int64_t)Fp->m[0] * Fp->m[1],
                                               dp<double>(cv), 0, Fp->m[0], (int64_t)Fp->m[0] * Fp->m[1], nullptr, s));
                    HIP_OK(c, hipMemcpyAsync(xh.data(), cv.get(), sizeof(double) * n, hipMemcpyDeviceToHost, s));
                    HIP_OK(c, hipStreamSynchronize(s));
                }
                double lam = 0.0;
                for (double v : xh) lam += v * v;
                lam = std::sqrt(lam);
                if (!(lam > 0.0)) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: zero power iterate");
                q.tau = 1.0 / lam;
                bool keep = l > nl - 5;
                for (int j = 0; j < cd.B.d; ++j) {
                    int nel = 0;
                    for (size_t t = 0; t + 1 < cd.B.knots[j].size(); ++t) nel += cd.B.knots[j][t + 1] > cd.B.knots[j][t];
                    keep = keep && nel > cd.B.p[j];
                }
                if (keep) saved = q.tau;
            } else {
                q.tau = saved;
            }
        } else {
            if (n > 2048) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: a lower-dimensional "
                                                      "component has more than 2048 DOFs (dense solve)");
            for (int j = 0; j < n; ++j) pos[(size_t)cd.idx[(size_t)j]] = j;
            std::vector<double> Acc((size_t)n * n, 0.0), X((size_t)n * n, 0.0);
            for (int a = 0; a < n; ++a) {
                const int r = cd.idx[(size_t)a];
                for (int t = rp[(size_t)r]; t < rp[(size_t)r + 1]; ++t) {
                    const int b = pos[(size_t)ci[(size_t)t]];
                    if (b >= 0) Acc[(size_t)a * n + b] = va[(size_t)t];
                }
                X[(size_t)a * n + a] = 1.0;
            }
            for (int j = 0; j < n; ++j) pos[(size_t)cd.idx[(size_t)j]] = -1;
            if (!gsx::dense::lu_solve(n, Acc, X, n))
                return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: singular component matrix");
            q.Ainv = dalloc(c, sizeof(double) * (size_t)n * n);
            if (!q.Ainv) return fail(c, GSX_ERR_OOM, "device allocation (component inverse)");
            HIP_OK(c, hipMemcpyAsync(q.Ainv.get(), X.data(), sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice, s));
        }
        S->nmax = std::max(S->nmax, n);
        S->comps.push_back(std::move(q));
    }
    if (covered != V.n) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: components do not cover the DOFs");
    S->rc = dalloc(c, sizeof(double) * std::max(1, S->nmax));
    S->zc = dalloc(c, sizeof(double) * std::max(1, S->nmax));
    if (!S->rc || !S->zc) return fail(c, GSX_ERR_OOM, "device allocation (smoother)");
    HIP_OK(c, hipStreamSynchronize(s));
    V.split = S;
    V.scms.reset();
    return GSX_OK;
}

// Set-up of level l's smoother (1 <= l < nlevels).  GSX_SMOOTH_SCMS: the
// subspace corrected mass smoother of setupSubspaceCorrectedMassSmoother
// (examples/multiGrid_example.cpp:407-560) for one patch with every side
// Dirichlet (the interior is the only component with free functions): the
// smoother of the level's basis (sigma = MG.Scaling, alpha = 0, beta = 1)
// damped by 1 / lambda, lambda from ten power iterations of S A
// (gsPreconditionerOp::estimateLargestEigenvalueOfPreconditionedSystem,
// src/gsSolver/gsPreconditioner.h:121-132) started from v0_i = 1 / (1 + i
// mod 5); once a level above nlevels - 5 has more elements than the degree
// in every direction, its damping is kept (mg->scms_saved) for the finer
// levels, as the example's patchLocalDampingParameters are.
static gsx_status mg_setup_level_smoother(gsx_mg *mg, int l, int kind, double param, double damping)
{
    gsx_ctx *c = mg->ctx;
    hipStream_t s = c->stream;
    const int nl = (int)mg->L.size();
    MGLevel &V = mg->L[l];
    const int n = V.n;
    if (kind == GSX_SMOOTH_SCMS) {
        const double sigma = param > 0.0 ? param : 0.12;
        ScmsBox B;
        bool ok;
        if (l == nl - 1) {
            ok = mg->fine_box_ok;
            if (ok) B = mg->fine_box;
        } else {
            ok = scms_box_of(V.prob, B);
        }
        bool alldir = ok;
        for (int k = 0; ok && k < B.d; ++k) alldir = alldir && B.lo[k] == 1 && B.hi[k] == B.nb[k] - 2;
        if (!alldir) {
            // several components: patches, interfaces, sides without a
            // Dirichlet condition, edges, corners
            if (mg->precision == GSX_MIXED)
                return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother with several components: FP64 "
                                            "cycles only");
            if (l == nl - 1 ? !mg->fine_layout_ok : !V.prob)
                return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: needs a hierarchy built from a "
                                            "problem (gsx_mg_create or gsx_pmg_create)");
            GSX_TRY(scms_split_build(mg, l, l == nl - 1 ? mg->fine_layout : layout_of(V.prob), sigma));
            V.lmax = 0.0;
            V.smoother = kind;
            return GSX_OK;
        }
        std::unique_ptr<gsx_fd> F;
        gsx_status e = scms_build(c, B, sigma, 0.0, 1.0, F);
        if (e != GSX_OK) return e;
        if (F->size != V.n) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: level size mismatch");
        V.scms = std::shared_ptr<gsx_fd>(F.release());
        if (!V.d) V.d = dalloc(c, sizeof(double) * n);
        if (mg->precision == GSX_MIXED && !V.d32) V.d32 = dalloc(c, sizeof(float) * n);
        if (!V.d || (mg->precision == GSX_MIXED && !V.d32)) return fail(c, GSX_ERR_OOM, "device allocation (smoother)");
        if (mg->scms_saved == 0.0) {
            std::vector<double> xh(n);
            for (int i = 0; i < n; ++i) xh[i] = 1.0 / (1.0 + (double)(i % 5));
            DevPtr xv = dalloc(c, sizeof(double) * n), tv = dalloc(c, sizeof(double) * n);
            if (!xv || !tv) return fail(c, GSX_ERR_OOM, "device allocation (power iteration)");
            const gsx_fd *Fp = V.scms.get();
            for (int it = 0; it < 10; ++it) {
                double nrm = 0.0;
                for (int i = 0; i < n; ++i) nrm += xh[i] * xh[i];
                nrm = std::sqrt(nrm);
                for (int i = 0; i < n; ++i) xh[i] /= nrm;
                HIP_OK(c, hipMemcpyAsync(xv.get(), xh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
                HIP_OK(c, gsx::launch_spmv_f64(V.lpr, n, dp<int>(V.rowptr), dp<int>(V.col), dp<double>(V.val),
                                               dp<double>(xv), dp<double>(tv), nullptr, nullptr, 0, nullptr, s));
                HIP_OK(c, fd_apply<double>(Fp, dp<double>(tv), 0, Fp->m[0], (int64_t)Fp->m[0] * Fp->m[1],
                                           dp<double>(xv), 0, Fp->m[0], (int64_t)Fp->m[0] * Fp->m[1], nullptr, s));
                HIP_OK(c, hipMemcpyAsync(xh.data(), xv.get(), sizeof(double) * n, hipMemcpyDeviceToHost, s));
                HIP_OK(c, hipStreamSynchronize(s));
            }
            double lam = 0.0;
            for (int i = 0; i < n; ++i) lam += xh[i] * xh[i];
            lam = std::sqrt(lam);
            if (!(lam > 0.0)) return fail(c, GSX_ERR_ARG, "subspace corrected mass smoother: zero power iterate");
            V.scms_tau = 1.0 / lam;
            // elements per direction: the distinct non-empty knot spans
            // (gsTensorBSplineBasis::component(j).numElements())
            bool keep = l > nl - 5;
            for (int k = 0; k < B.d; ++k) {
                int nel = 0;
                for (size_t q = 0; q + 1 < B.knots[k].size(); ++q) nel += B.knots[k][q + 1] > B.knots[k][q];
                keep = keep && nel > B.p[k];
            }
            if (keep) mg->scms_saved = V.scms_tau;
        } else {
            V.scms_tau = mg->scms_saved;
        }
        V.lmax = V.scms_tau;
        V.smoother = kind;
        return GSX_OK;
    }
    V.smoother = kind;
    V.tau = damping;
    if (kind == GSX_SMOOTH_JACOBI) return GSX_OK;
    if (kind == GSX_SMOOTH_GAUSS_SEIDEL) {
        if (V.gs_fwd) return GSX_OK;
        std::vector<int> rp(V.n + 1), ci((size_t)V.nnz), rows;
        HIP_OK(c, hipMemcpyAsync(rp.data(), V.rowptr.get(), sizeof(int) * (V.n + 1), hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipMemcpyAsync(ci.data(), V.col.get(), sizeof(int) * V.nnz, hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipStreamSynchronize(s));
        V.gs_fwd = dalloc(c, sizeof(int) * std::max(1, V.n));
        V.gs_rev = dalloc(c, sizeof(int) * std::max(1, V.n));
        if (!V.gs_fwd || !V.gs_rev) return fail(c, GSX_ERR_OOM, "device allocation (Gauss-Seidel schedule)");
        gs_schedule(V.n, rp, ci, false, rows, V.gs_fwd_off);
        HIP_OK(c, hipMemcpyAsync(V.gs_fwd.get(), rows.data(), sizeof(int) * V.n, hipMemcpyHostToDevice, s));
        HIP_OK(c, hipStreamSynchronize(s));
        gs_schedule(V.n, rp, ci, true, rows, V.gs_rev_off);
        HIP_OK(c, hipMemcpyAsync(V.gs_rev.get(), rows.data(), sizeof(int) * V.n, hipMemcpyHostToDevice, s));
        HIP_OK(c, hipStreamSynchronize(s));
        return GSX_OK;
    }
    if (kind == GSX_SMOOTH_CHEBYSHEV) V.cheb_ratio = param > 1.0 ? param : 30.0;
    if (!V.d) V.d = dalloc(c, sizeof(double) * n);
    if (mg->precision == GSX_MIXED && !V.d32) V.d32 = dalloc(c, sizeof(float) * n);
    if (!V.d || (mg->precision == GSX_MIXED && !V.d32)) return fail(c, GSX_ERR_OOM, "device allocation (smoother)");
    if (kind != GSX_SMOOTH_CHEBYSHEV) return GSX_OK;
    // lambda_max(D^-1 A): 20 power iterations from v0_i = 1 / (1 + i mod 5)
    std::vector<double> v0(n);
    double nv = 0.0;
    for (int i = 0; i < n; ++i) {
        v0[i] = 1.0 / (1.0 + (double)(i % 5));
        nv += v0[i] * v0[i];
    }
    nv = std::sqrt(nv);
    for (int i = 0; i < n; ++i) v0[i] /= nv;
    const int nb = (n + 255) / 256;
    DevPtr v = dalloc(c, sizeof(double) * n), w = dalloc(c, sizeof(double) * n), part = dalloc(c, sizeof(double) * nb);
    if (!v || !w || !part) return fail(c, GSX_ERR_OOM, "device allocation (power iteration)");
    HIP_OK(c, hipMemcpyAsync(v.get(), v0.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    std::vector<double> ph(nb);
    double lam = 0.0;
    for (int it = 0; it < 20; ++it) {
        HIP_OK(c, gsx::launch_spmv_f64(V.lpr, n, dp<int>(V.rowptr), dp<int>(V.col), dp<double>(V.val), dp<double>(v),
                                       dp<double>(w), nullptr, nullptr, 0, nullptr, s));
        HIP_OK(c, gsx::launch_div_sq(n, dp<double>(w), dp<double>(V.diag), dp<double>(part), s));
        HIP_OK(c, hipMemcpyAsync(ph.data(), part.get(), sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipStreamSynchronize(s));
        double sq = 0.0;
        for (int k = 0; k < nb; ++k) sq += ph[k];
        lam = std::sqrt(sq);
        HIP_OK(c, gsx::launch_axpby(n, 1.0 / lam, dp<double>(w), 0.0, dp<double>(v), s));
    }
    V.lmax = 1.1 * lam;
    HIP_OK(c, hipStreamSynchronize(s));
    return GSX_OK;
}

extern "C" gsx_status gsx_mg_set_smoother(gsx_mg *mg, int32_t kind, double param)
{
    if (!mg || kind < GSX_SMOOTH_JACOBI || kind > GSX_SMOOTH_SCMS) return GSX_ERR_ARG;
    mg->scms_saved = 0.0;
    for (int l = 1; l < (int)mg->L.size(); ++l) GSX_TRY(mg_setup_level_smoother(mg, l, kind, param, mg->tau));
    return GSX_OK;
}

extern "C" gsx_status gsx_mg_set_cycle(gsx_mg *mg, int32_t npre, int32_t npost, double cgc_damping)
{
    if (!mg || npre < 0 || npost < 0) return GSX_ERR_ARG;
    mg->npre = npre;
    mg->npost = npost;
    mg->cgc = cgc_damping;
    return GSX_OK;
}

extern "C" gsx_status gsx_mg_set_level_smoother(gsx_mg *mg, int32_t lvl, int32_t kind, double damping, double param)
{
    if (!mg || kind < GSX_SMOOTH_JACOBI || kind > GSX_SMOOTH_SCMS) return GSX_ERR_ARG;
    if (lvl < 1 || lvl >= (int)mg->L.size()) return fail(mg->ctx, GSX_ERR_ARG, "smoother level out of range");
    if (lvl == 1) mg->scms_saved = 0.0;
    return mg_setup_level_smoother(mg, lvl, kind, param, damping);
}

extern "C" gsx_status gsx_mg_smoother_bound(const gsx_mg *mg, int32_t lvl, double *lmax)
{
    if (!mg || !lmax || lvl < 0 || lvl >= (int)mg->L.size()) return GSX_ERR_ARG;
    *lmax = mg->L[lvl].lmax;
    return GSX_OK;
}

extern "C" gsx_status gsx_mg_apply(gsx_mg *mg, const double *r_dev, double *z_dev)
{
    if (!mg || !r_dev || !z_dev) return GSX_ERR_ARG;
    return mg_apply_dev(mg, r_dev, z_dev);
}

// MG-preconditioned CG (FP64 outer iteration on A; the cycle in the
// precision of mg).  Host-controlled: one state read per iteration.
static gsx_status cg_mg(gsx_matrix *A, gsx_mg *mg, const double *b, double *x, double tol, int maxit,
                        int *iters, double *relres)
{
    gsx_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    gsx_status e = cg_workspace(A, std::max(1, maxit), false, false);
    if (e != GSX_OK) return e;
    CGWork &w = A->ws;
    const int n = A->n, np = w.nparts;
    const int ns = gsx::spmv_blocks(n, A->lpr), nv = (n + 255) / 256;
    double *P0 = dp<double>(w.parts), *P1 = P0 + np;
    CGState *st = dp<CGState>(w.state);
    double *hist = dp<double>(w.hist);
    const int *rp = dp<int>(A->rowptr), *ci = dp<int>(A->col);
    const double *va = dp<double>(A->val);
    double *r = dp<double>(w.r), *z = dp<double>(w.z), *p = dp<double>(w.p), *t = dp<double>(w.t);
    CGState h{};
    h.tol = tol;
    h.maxit = maxit;
    HIP_OK(c, hipMemcpyAsync(st, &h, sizeof(CGState), hipMemcpyHostToDevice, s));
    HIP_OK(c, gsx::launch_spmv_f64(A->lpr, n, rp, ci, va, x, r, b, P1, 2, nullptr, s));
    HIP_OK(c, gsx::launch_dot_f64(n, b, b, P0, nullptr, s));
    HIP_OK(c, gsx::launch_cg_scalar(0, P0, nv, P1, ns, st, hist, s));
    GSX_TRY(read_state(c, w, h));
    if (h.done == 3) HIP_OK(c, gsx::launch_fill_f64(x, n, 0.0, s));
    if (!h.done) {
        GSX_TRY(mg_apply_dev(mg, r, z));
        HIP_OK(c, hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        HIP_OK(c, gsx::launch_dot_f64(n, r, z, P0, nullptr, s));
        HIP_OK(c, gsx::launch_cg_scalar(1, P0, nv, nullptr, 0, st, hist, s));
        while (true) {
            {
                KScope k(c, "spmv_f64");
                HIP_OK(c, gsx::launch_spmv_f64(A->lpr, n, rp, ci, va, p, t, nullptr, P0, 1, st, s));
            }
            HIP_OK(c, gsx::launch_cg_scalar(2, P0, ns, nullptr, 0, st, hist, s));
            HIP_OK(c, gsx::launch_cg_xr_f64(n, x, r, p, t, P1, st, s));
            HIP_OK(c, gsx::launch_cg_scalar(5, P1, nv, nullptr, 0, st, hist, s));
            GSX_TRY(read_state(c, w, h));
            if (h.done) break;
            GSX_TRY(mg_apply_dev(mg, r, z));
            HIP_OK(c, gsx::launch_dot_f64(n, r, z, P0, nullptr, s));
            HIP_OK(c, gsx::launch_cg_scalar(6, P0, nv, nullptr, 0, st, hist, s));
            HIP_OK(c, gsx::launch_cg_pupdate_f64(n, p, z, st, s));
        }
    }
    *iters = h.it;
    *relres = h.err;
    A->hist.resize(h.it + 1);
    HIP_OK(c, hipMemcpyAsync(A->hist.data(), hist, sizeof(double) * (h.it + 1), hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    if (h.done == 1 || h.done == 3 || h.err < tol) return GSX_OK;
    return fail(c, GSX_ERR_NOT_CONVERGED, "MG-PCG reached the maximum number of iterations");
}

static gsx_status cg_mg_entry(gsx_matrix *A, gsx_mg *mg, const double *b, double *x, double tol, int maxit,
                              int *iters, double *relres)
{
    return cg_mg(A, mg, b, x, tol, maxit, iters, relres);
}

// ===========================================================================
// Batched CG: independent solves advanced together (gsx_batch.hip)
// ===========================================================================
namespace {

struct BMember {
    bool used = false;
    int status = 0;           // 0 running, else gsx::BState::done
    gsx_matrix *A = nullptr;
    int row0 = 0, nrows = 0;
    int maxit = 0;
    DevPtr r, z, p, t, r32, z32, p32, t32, xc32, d32, hist;
    DevPtr s64, s32;          // structured operator of a box member
    gsx::BSeg seg{};
    gsx::BState st{};
};

template <typename T>
T *shift(T *p, int row0)
{
    return p ? reinterpret_cast<T *>(reinterpret_cast<uintptr_t>(p) - sizeof(T) * (size_t)row0) : nullptr;
}

}  // namespace

struct gsx_batch {
    gsx_ctx *c = nullptr;
    bool mixed = false;
    int cap = 0;
    std::vector<BMember> m;
    DevPtr d_segs, d_state, d_chunks, d_parts, d_bad, d_groups, d_schunks;
    int chunk_cap = 0, nchunks = 0, nschunks = 0, nsingle = 0;
    int group_max = 1;                   // members per SpMV group (1: no groups)
    int group_rows = 131072;             // rows from which a member joins a group
    bool dirty = true;
    std::vector<gsx::BChunk> chunks_h;
    std::vector<gsx::BGroup> groups_h;   // members sharing an operator and tile shape
    std::vector<DevPtr> group_bufs;       // their interleaved SpMV inputs
    std::vector<gsx::BSChunk> schunks_h; // the SpMV launch's chunks (one per group chunk)
    std::vector<gsx::BState> state_h;
    // one step (init, iters iterations, reliable update) captured as a HIP
    // graph for the current member layout; replayed while it holds
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_iters = -1;
    int64_t graph_launches = 0;
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        graph_iters = -1;
    }
    ~gsx_batch() { drop_graph(); }
};

gsx_status gsx_batch_create(gsx_ctx *c, int32_t precision, int32_t capacity, gsx_batch **out)
{
    if (!c || !out || capacity <= 0 || (precision != GSX_FP64 && precision != GSX_MIXED)) return GSX_ERR_ARG;
    gsx_batch *B = new gsx_batch();
    B->c = c;
    B->mixed = precision == GSX_MIXED;
    B->cap = capacity;
    B->m.resize(capacity);
    B->state_h.resize(capacity);
    B->d_segs = dalloc(c, sizeof(gsx::BSeg) * capacity);
    B->d_groups = dalloc(c, sizeof(gsx::BGroup) * capacity);
    B->d_state = dalloc(c, sizeof(gsx::BState) * capacity);
    B->d_bad = dalloc(c, sizeof(int));
    if (!B->d_segs || !B->d_state || !B->d_bad || !B->d_groups) {
        delete B;
        return fail(c, GSX_ERR_OOM, "device allocation (batch)");
    }
    *out = B;
    return GSX_OK;
}

gsx_status gsx_batch_destroy(gsx_batch *B)
{
    if (!B) return GSX_ERR_ARG;
    (void)hipStreamSynchronize(B->c->stream);
    delete B;
    return GSX_OK;
}

static gsx_status batch_add(gsx_batch *B, gsx_matrix *A, int32_t row0, int32_t nrows, const int32_t *box,
                            int32_t precond, double tau, const double *b, double *x, double tol, int32_t maxit,
                            int32_t *id)
{
    if (!B || !A || !b || !x || !id || maxit < 1 || nrows <= 0 || row0 < 0 || (int64_t)row0 + nrows > A->n)
        return GSX_ERR_ARG;
    gsx_ctx *c = B->c;
    if (box) {
        if ((box[5] != 1 && box[5] != 2 && box[5] != 4) || (box[6] != 0 && box[6] != 1))
            return fail(c, GSX_ERR_ARG, "box member: lanes 1, 2 or 4 and half 0 or 1");
        const int d = box[0], p = box[1];
        const int64_t cells = (int64_t)box[2] * box[3] * (d == 3 ? box[4] : 1);
        if ((d != 2 && d != 3) || p < 1 || p > 4 || box[2] < 1 || box[3] < 1 || (d == 3 && box[4] < 1) ||
            (d == 2 && box[4] != 1) || cells != nrows)
            return fail(c, GSX_ERR_ARG, "box member: d in {2, 3}, p in 1..4, m0 m1 m2 = nrows (m2 = 1 in 2D)");
    }
    if (A->ctx->device != c->device) return fail(c, GSX_ERR_ARG, "matrix lives on another device");
    if (precond != GSX_PREC_NONE && precond != GSX_PREC_JACOBI)
        return fail(c, GSX_ERR_ARG, "batched CG supports no preconditioner or Jacobi");
    if (B->mixed && !A->mixed) return fail(c, GSX_ERR_ARG, "mixed batch needs matrices assembled with GSX_MIXED");
    int slot = -1;
    for (int k = 0; k < B->cap; ++k)
        if (!B->m[k].used) {
            slot = k;
            break;
        }
    if (slot < 0) return fail(c, GSX_ERR_ARG, "batch is full");
    hipStream_t s = c->stream;
    // A box member that is the whole of a single-patch matrix whose
    // structured operator the assembly has filled (the same box) iterates on
    // that operator itself: no copy, no checks (the rows are the patch's)
    const StencilOp &S = A->st;
    const bool own = box && row0 == 0 && nrows == A->n && S.available && S.s64_ready && S.d == box[0] &&
                     S.p == box[1] && S.m[0] == box[2] && S.m[1] == box[3] && (S.d == 2 || S.m[2] == box[4]) &&
                     (!B->mixed || S.s32_ready);
    int rp2[2] = {0, 0};
    int bad = 0;
    if (!own) {
        // the member's rows must only couple to its own rows
        HIP_OK(c, hipMemcpyAsync(&rp2[0], dp<int>(A->rowptr) + row0, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipMemcpyAsync(&rp2[1], dp<int>(A->rowptr) + row0 + nrows, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipMemsetAsync(B->d_bad.get(), 0, sizeof(int), s));
        HIP_OK(c, gsx::launch_batch_check(row0, nrows, dp<int>(A->rowptr), dp<int>(A->col), dp<int>(B->d_bad), s));
        HIP_OK(c, hipMemcpyAsync(&bad, B->d_bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipStreamSynchronize(s));
        if (bad) return fail(c, GSX_ERR_ARG, "member rows couple to rows outside [row0, row0 + nrows)");
    }
    if (!A->diag && precond == GSX_PREC_JACOBI) {
        A->diag = dalloc(c, sizeof(double) * A->n);
        if (!A->diag) return fail(c, GSX_ERR_OOM, "device allocation (diag)");
        HIP_OK(c, gsx::launch_diag_f64(A->n, dp<int>(A->rowptr), dp<int>(A->col), dp<double>(A->val),
                                       dp<double>(A->diag), s));
    }
    if (B->mixed && !A->val32) {
        A->val32 = dalloc(c, sizeof(float) * std::max<int64_t>(1, A->nnz));
        if (!A->val32) return fail(c, GSX_ERR_OOM, "device allocation (fp32 values)");
        HIP_OK(c, gsx::launch_f64_to_f32(A->nnz, dp<double>(A->val), dp<float>(A->val32), s));
    }
    BMember M;
    M.used = true;
    M.A = A;
    M.row0 = row0;
    M.nrows = nrows;
    M.maxit = maxit;
    const size_t n8 = sizeof(double) * nrows, n4 = sizeof(float) * nrows;
    M.r = dalloc(c, n8);
    M.z = dalloc(c, n8);
    M.p = dalloc(c, n8);
    M.t = dalloc(c, n8);
    M.hist = dalloc(c, sizeof(double) * (maxit + 1));
    bool ok = M.r && M.z && M.p && M.t && M.hist;
    if (B->mixed) {
        M.r32 = dalloc(c, n4);
        M.z32 = dalloc(c, n4);
        M.p32 = dalloc(c, n4);
        M.t32 = dalloc(c, n4);
        M.xc32 = dalloc(c, n4);
        ok = ok && M.r32 && M.z32 && M.p32 && M.t32 && M.xc32;
        if (precond == GSX_PREC_JACOBI) {
            M.d32 = dalloc(c, n4);
            ok = ok && M.d32;
            if (ok) HIP_OK(c, gsx::launch_f64_to_f32(nrows, dp<double>(A->diag) + row0, dp<float>(M.d32), s));
        }
    }
    if (!ok) return fail(c, GSX_ERR_OOM, "device allocation (batch member)");
    gsx::BSeg &g = M.seg;
    g.rp = dp<int>(A->rowptr);
    g.ci = dp<int>(A->col);
    g.v64 = dp<double>(A->val);
    g.v32 = B->mixed ? dp<float>(A->val32) : nullptr;
    g.b = shift(b, row0);
    g.x = shift(x, row0);
    g.r = shift(dp<double>(M.r), row0);
    g.z = shift(dp<double>(M.z), row0);
    g.p = shift(dp<double>(M.p), row0);
    g.t = shift(dp<double>(M.t), row0);
    g.r32 = shift(dp<float>(M.r32), row0);
    g.z32 = shift(dp<float>(M.z32), row0);
    g.p32 = shift(dp<float>(M.p32), row0);
    g.t32 = shift(dp<float>(M.t32), row0);
    g.xc32 = shift(dp<float>(M.xc32), row0);
    g.d64 = precond == GSX_PREC_JACOBI ? dp<double>(A->diag) : nullptr;
    g.d32 = shift(dp<float>(M.d32), row0);
    g.hist = dp<double>(M.hist);
    g.tau = tau;
    g.tol = tol;
    g.maxit = maxit;
    g.lpr = own ? 8 : gsx::spmv_lpr((double)(rp2[1] - rp2[0]) / nrows);
    g.row0 = row0;
    if (own) {
        const int d = box[0], p = box[1];
        g.s64 = dp<double>(S.s64);
        g.s32 = B->mixed ? dp<float>(S.s32) : nullptr;
        if (B->mixed && box[6] == 0) {
            // the full FP32 form asked for: one copy from the FP64 slots,
            // kept with the matrix for every member on it
            StencilOp &So = A->st;
            if (!So.s32full) {
                const int64_t ns = gsx::batch_box_slots(d, p, false);
                So.s32full = dalloc(c, sizeof(float) * gsx::stencil_size(nrows, (int)ns));
                if (!So.s32full) return fail(c, GSX_ERR_OOM, "device allocation (box member operator)");
                HIP_OK(c, gsx::launch_batch_box_s32(d, p, nrows, false, dp<double>(S.s64), dp<float>(So.s32full), s));
            }
            g.s32 = dp<float>(So.s32full);
        }
        g.bd = d;
        g.bp = p;
        g.tpr = box[5];
        g.half = box[6];
        for (int k = 0; k < 3; ++k) g.bm[k] = box[2 + k];
    } else if (box) {
        // the member's structured operator, built from its CSR rows (which
        // must lie in the box's stencil window)
        const int d = box[0], p = box[1];
        const bool half = box[6] != 0;
        const int64_t n64 = gsx::stencil_size(nrows, (int)gsx::batch_box_slots(d, p, false));
        const int64_t n32 = gsx::stencil_size(nrows, (int)gsx::batch_box_slots(d, p, half));
        M.s64 = dalloc(c, sizeof(double) * n64);
        if (B->mixed) M.s32 = dalloc(c, sizeof(float) * n32);
        if (!M.s64 || (B->mixed && !M.s32)) return fail(c, GSX_ERR_OOM, "device allocation (box member operator)");
        KScope ks(c, "batch_box_build");
        HIP_OK(c, gsx::launch_fill_f64(dp<double>(M.s64), n64, 0.0, s));
        if (B->mixed) HIP_OK(c, gsx::launch_fill_f32(dp<float>(M.s32), n32, 0.0f, s));
        HIP_OK(c, hipMemsetAsync(B->d_bad.get(), 0, sizeof(int), s));
        HIP_OK(c, gsx::launch_batch_box_build(d, p, row0, nrows, box[2], box[3], half, dp<int>(A->rowptr),
                                              dp<int>(A->col), dp<double>(A->val), dp<double>(M.s64),
                                              B->mixed ? dp<float>(M.s32) : nullptr, dp<int>(B->d_bad), s));
        HIP_OK(c, hipMemcpyAsync(&bad, B->d_bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipStreamSynchronize(s));
        if (bad) return fail(c, GSX_ERR_ARG, "box member: an entry lies outside the box's stencil window");
        g.bd = d;
        g.bp = p;
        g.tpr = box[5];
        g.half = half ? 1 : 0;
        for (int k = 0; k < 3; ++k) g.bm[k] = box[2 + k];
        g.s64 = dp<double>(M.s64);
        g.s32 = dp<float>(M.s32);
    }
    std::memset(&M.st, 0, sizeof(M.st));
    // fresh state: phase 0 (initialised by the next step)
    HIP_OK(c, hipMemcpyAsync(dp<gsx::BState>(B->d_state) + slot, &M.st, sizeof(gsx::BState),
                             hipMemcpyHostToDevice, s));
    HIP_OK(c, hipStreamSynchronize(s));
    B->m[slot] = std::move(M);
    B->dirty = true;
    *id = slot;
    return GSX_OK;
}

gsx_status gsx_batch_add(gsx_batch *B, gsx_matrix *A, int32_t row0, int32_t nrows, int32_t precond, double tau,
                         const double *b, double *x, double tol, int32_t maxit, int32_t *id)
{
    return batch_add(B, A, row0, nrows, nullptr, precond, tau, b, x, tol, maxit, id);
}

gsx_status gsx_batch_add_box(gsx_batch *B, gsx_matrix *A, int32_t row0, int32_t nrows, int32_t d, int32_t p,
                             int32_t m0, int32_t m1, int32_t m2, int32_t lanes, int32_t half, int32_t precond,
                             double tau, const double *b, double *x, double tol, int32_t maxit, int32_t *id)
{
    const int32_t box[7] = {d, p, m0, m1, m2, lanes, half};
    return batch_add(B, A, row0, nrows, box, precond, tau, b, x, tol, maxit, id);
}

gsx_status gsx_batch_set_tile(gsx_batch *B, int32_t id, int32_t lanes, int32_t half)
{
    if (!B || id < 0 || id >= B->cap || !B->m[id].used) return GSX_ERR_ARG;
    gsx_ctx *c = B->c;
    BMember &M = B->m[id];
    gsx::BSeg &g = M.seg;
    if (M.st.phase != 0) return fail(c, GSX_ERR_ARG, "tile shape: set it before the member's first step");
    if (!g.bd) {
        if (lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
            return fail(c, GSX_ERR_ARG, "tile shape: CSR members take 8, 16, 32 or 64 lanes per row");
        g.lpr = lanes;
    } else {
        if (lanes != 1 && lanes != 2 && lanes != 4)
            return fail(c, GSX_ERR_ARG, "tile shape: box members take 1, 2 or 4 threads per row");
        if (half != 0 && half != 1) return GSX_ERR_ARG;
        if (B->mixed && half != g.half) {
            // the FP32 copy in the other form, from the FP64 stencil
            HIP_OK(c, hipStreamSynchronize(c->stream));   // kernels may still read the old copy
            const int64_t ns = gsx::batch_box_slots(g.bd, g.bp, half != 0);
            DevPtr s32 = dalloc(c, sizeof(float) * gsx::stencil_size(M.nrows, (int)ns));
            if (!s32) return fail(c, GSX_ERR_OOM, "device allocation (box member operator)");
            HIP_OK(c, gsx::launch_batch_box_s32(g.bd, g.bp, M.nrows, half != 0, dp<double>(M.s64), dp<float>(s32),
                                                c->stream));
            M.s32 = std::move(s32);
            g.s32 = dp<float>(M.s32);
        }
        g.tpr = lanes;
        g.half = half;
    }
    B->dirty = true;
    return GSX_OK;
}

static gsx_status batch_rebuild(gsx_batch *B)
{
    gsx_ctx *c = B->c;
    hipStream_t s = c->stream;
    B->chunks_h.clear();
    for (int k = 0; k < B->cap; ++k) {
        BMember &M = B->m[k];
        if (!M.used || M.status) continue;
        // box members: one stencil tile of the member per chunk
        const int rows = M.seg.bd ? gsx::kStencilTile / M.seg.tpr : gsx::batch_chunk_rows(M.seg.lpr);
        M.seg.c0 = (int)B->chunks_h.size();
        for (int lo = M.row0; lo < M.row0 + M.nrows; lo += rows)
            B->chunks_h.push_back({k, lo, std::min(lo + rows, M.row0 + M.nrows)});
        M.seg.nc = (int)B->chunks_h.size() - M.seg.c0;
        HIP_OK(c, hipMemcpyAsync(dp<gsx::BSeg>(B->d_segs) + k, &M.seg, sizeof(gsx::BSeg), hipMemcpyHostToDevice, s));
    }
    B->nchunks = (int)B->chunks_h.size();
    // SpMV groups: box members on one operator with one tile shape (up to
    // kGroupMax), every other member alone
    B->groups_h.clear();
    B->schunks_h.clear();
    B->nsingle = 0;
    std::vector<int> gfirst;   // first member of each group
    for (int k = 0; k < B->cap; ++k) {
        const BMember &M = B->m[k];
        if (!M.used || M.status) continue;
        const gsx::BSeg &g = M.seg;
        int gid = -1;
        if (g.bd && M.nrows >= B->group_rows)
            for (size_t q = 0; q < B->groups_h.size(); ++q) {
                const gsx::BSeg &h = B->m[gfirst[q]].seg;
                if (B->groups_h[q].k < B->group_max && h.bd && h.s64 == g.s64 && h.s32 == g.s32 &&
                    h.tpr == g.tpr && h.half == g.half && h.row0 == g.row0 && h.nc == g.nc) {
                    gid = (int)q;
                    break;
                }
            }
        if (gid < 0) {
            gid = (int)B->groups_h.size();
            B->groups_h.push_back(gsx::BGroup{});
            gfirst.push_back(k);
        }
        gsx::BGroup &G = B->groups_h[gid];
        G.seg[G.k++] = k;
    }
    // interleaved SpMV inputs of the groups of several members (kGroupMax
    // values per row of their box)
    B->group_bufs.clear();
    for (size_t q = 0; q < B->groups_h.size(); ++q) {
        gsx::BGroup &G = B->groups_h[q];
        G.xi = nullptr;
        G.xrow0 = 0;
        if (G.k < 2) continue;
        const BMember &M = B->m[gfirst[q]];
        DevPtr buf = dalloc(c, (B->mixed ? sizeof(float) : sizeof(double)) * (size_t)M.nrows * gsx::kGroupMax);
        if (!buf) return fail(c, GSX_ERR_OOM, "device allocation (group inputs)");
        G.xi = buf.get();
        G.xrow0 = M.row0;
        B->group_bufs.push_back(std::move(buf));
    }
    // chunks of the single-member groups first, then those of the others
    // (two launches: the single-member kernel keeps its register budget)
    for (int pass = 0; pass < 2; ++pass)
        for (size_t q = 0; q < B->groups_h.size(); ++q) {
            if ((B->groups_h[q].k > 1) != (pass == 1)) continue;
            const gsx::BSeg &h = B->m[gfirst[q]].seg;
            for (int i = 0; i < h.nc; ++i) {
                const gsx::BChunk &ch = B->chunks_h[h.c0 + i];
                B->schunks_h.push_back({(int)q, ch.lo, ch.hi, i});
            }
            if (pass == 0) B->nsingle = (int)B->schunks_h.size();
        }
    B->nschunks = (int)B->schunks_h.size();
    if (B->nchunks > B->chunk_cap) {
        HIP_OK(c, hipStreamSynchronize(s));
        B->chunk_cap = std::max(B->nchunks, 2 * B->chunk_cap);
        B->d_chunks = dalloc(c, sizeof(gsx::BChunk) * B->chunk_cap);
        B->d_schunks = dalloc(c, sizeof(gsx::BSChunk) * B->chunk_cap);
        B->d_parts = dalloc(c, sizeof(double) * 3 * B->chunk_cap);
        if (!B->d_chunks || !B->d_schunks || !B->d_parts)
            return fail(c, GSX_ERR_OOM, "device allocation (batch chunks)");
    }
    if (!B->groups_h.empty())
        HIP_OK(c, hipMemcpyAsync(B->d_groups.get(), B->groups_h.data(), sizeof(gsx::BGroup) * B->groups_h.size(),
                                 hipMemcpyHostToDevice, s));
    if (B->nschunks)
        HIP_OK(c, hipMemcpyAsync(B->d_schunks.get(), B->schunks_h.data(), sizeof(gsx::BSChunk) * B->nschunks,
                                 hipMemcpyHostToDevice, s));
    if (B->nchunks)
        HIP_OK(c, hipMemcpyAsync(B->d_chunks.get(), B->chunks_h.data(), sizeof(gsx::BChunk) * B->nchunks,
                                 hipMemcpyHostToDevice, s));
    // the host arrays are reused by the next rebuild: finish the copies
    HIP_OK(c, hipStreamSynchronize(s));
    B->dirty = false;
    return GSX_OK;
}

gsx_status gsx_batch_step(gsx_batch *B, int32_t iters, int32_t *running)
{
    if (!B || iters < 0) return GSX_ERR_ARG;
    gsx_ctx *c = B->c;
    hipStream_t s = c->stream;
    if (B->dirty) {
        B->drop_graph();
        GSX_TRY(batch_rebuild(B));
    }
    const int nc = B->nchunks;
    if (nc > 0) {
        const gsx::BChunk *ch = dp<gsx::BChunk>(B->d_chunks);
        const gsx::BSeg *sg = dp<gsx::BSeg>(B->d_segs);
        gsx::BState *st = dp<gsx::BState>(B->d_state);
        double *P0 = dp<double>(B->d_parts), *P1 = P0 + B->chunk_cap, *P2 = P1 + B->chunk_cap;
        auto enqueue = [&]() -> hipError_t {
            hipError_t e = gsx::launch_batch_init(nc, ch, sg, st, P0, P1, P2, B->mixed, s);
            for (int i = 0; e == hipSuccess && i < iters; ++i)
                e = gsx::launch_batch_iteration(nc, ch, B->nsingle, B->nschunks - B->nsingle,
                                                dp<gsx::BSChunk>(B->d_schunks), dp<gsx::BGroup>(B->d_groups), sg,
                                                st, P0, P1, P2, B->mixed, 1e-4, s);
            if (e == hipSuccess && B->mixed) e = gsx::launch_batch_reliable(nc, ch, sg, st, P0, P2, s);
            return e;
        };
        if (B->use_graph && !c->timing) {
            // the step's launches depend only on the member layout and iters:
            // capture them once, replay every step until the layout changes
            if (!B->gexec || B->graph_iters != iters) {
                B->drop_graph();
                HIP_OK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                const hipError_t e = enqueue();
                hipGraph_t g = nullptr;
                const hipError_t e2 = hipStreamEndCapture(s, &g);
                if (e != hipSuccess || e2 != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    return fail(c, GSX_ERR_HIP, std::string("batch graph capture: ") +
                                                    hipGetErrorString(e != hipSuccess ? e : e2));
                }
                B->graph = g;
                HIP_OK(c, hipGraphInstantiate(&B->gexec, B->graph, nullptr, nullptr, 0));
                B->graph_iters = iters;
            }
            KScope k(c, "batch_graph");
            HIP_OK(c, hipGraphLaunch(B->gexec, s));
            ++B->graph_launches;
        } else {
            KScope k(c, B->mixed ? "batch_step_f32" : "batch_step_f64");
            HIP_OK(c, enqueue());
        }
        HIP_OK(c, hipMemcpyAsync(B->state_h.data(), st, sizeof(gsx::BState) * B->cap, hipMemcpyDeviceToHost, s));
        HIP_OK(c, hipStreamSynchronize(s));
    }
    int run = 0;
    for (int k = 0; k < B->cap; ++k) {
        BMember &M = B->m[k];
        if (!M.used || M.status) continue;
        M.st = B->state_h[k];
        if (M.st.phase == 1 && M.st.done) {
            M.status = M.st.done;
            B->dirty = true;
        } else {
            ++run;
        }
    }
    if (running) *running = run;
    return GSX_OK;
}

gsx_status gsx_batch_set_group(gsx_batch *B, int32_t group_max)
{
    if (!B || group_max < 1 || group_max > gsx::kGroupMax) return GSX_ERR_ARG;
    B->group_max = group_max;
    B->dirty = true;
    return GSX_OK;
}

gsx_status gsx_batch_set_group_rows(gsx_batch *B, int64_t rows)
{
    if (!B || rows < 1) return GSX_ERR_ARG;
    B->group_rows = (int)std::min<int64_t>(rows, INT32_MAX);
    B->dirty = true;
    return GSX_OK;
}

gsx_status gsx_batch_set_graph(gsx_batch *B, int32_t enable)
{
    if (!B) return GSX_ERR_ARG;
    B->use_graph = enable != 0;
    if (!B->use_graph) B->drop_graph();
    return GSX_OK;
}

gsx_status gsx_batch_graph_launches(const gsx_batch *B, int64_t *count)
{
    if (!B || !count) return GSX_ERR_ARG;
    *count = B->graph_launches;
    return GSX_OK;
}

gsx_status gsx_batch_result(const gsx_batch *B, int32_t id, int32_t *status, int32_t *iters, double *relres)
{
    if (!B || id < 0 || id >= B->cap || !B->m[id].used) return GSX_ERR_ARG;
    const BMember &M = B->m[id];
    if (status) *status = M.status;
    if (iters) *iters = M.st.it_next;
    if (relres) *relres = M.st.err;
    return GSX_OK;
}

gsx_status gsx_batch_history(const gsx_batch *B, int32_t id, double *hist, int32_t cap, int32_t *count)
{
    if (!B || id < 0 || id >= B->cap || !B->m[id].used || !count) return GSX_ERR_ARG;
    const BMember &M = B->m[id];
    const int n = std::min(M.st.it_next + 1, M.maxit + 1);
    *count = n;
    if (hist && cap > 0) {
        gsx_ctx *c = B->c;
        HIP_OK(c, hipMemcpyAsync(hist, M.hist.get(), sizeof(double) * std::min(n, cap), hipMemcpyDeviceToHost,
                                 c->stream));
        HIP_OK(c, hipStreamSynchronize(c->stream));
    }
    return GSX_OK;
}

gsx_status gsx_batch_remove(gsx_batch *B, int32_t id)
{
    if (!B || id < 0 || id >= B->cap || !B->m[id].used) return GSX_ERR_ARG;
    // kernels that may still read the member's buffers
    HIP_OK(B->c, hipStreamSynchronize(B->c->stream));
    B->m[id] = BMember();
    B->dirty = true;
    return GSX_OK;
}

// ===========================================================================
// Boundary data: Dirichlet values by L2 projection, Neumann load
// ===========================================================================
namespace {

// side elements of (patch, side) in the order of the side's domain iterator
// (lexicographic over the tangential directions, direction 0 fastest)
void side_items(const HostPatch &h, int patch, const gsx_bc &b, std::vector<gsx::BdryItem> &out)
{
    const int dir = (b.side - 1) / 2, par = (b.side - 1) % 2;
    int64_t cnt = 1;
    for (int k = 0; k < h.d; ++k)
        if (k != dir) cnt *= h.nel[k];
    for (int64_t t = 0; t < cnt; ++t) {
        gsx::BdryItem it{};
        it.patch = patch;
        it.side = b.side;
        it.type = b.type;
        it.kind = b.kind;
        for (int v = 0; v < 4; ++v) it.value[v] = b.value[v];
        int64_t r = t;
        for (int k = 0; k < h.d; ++k) {
            if (k == dir) {
                it.ei[k] = par ? h.nel[k] - 1 : 0;
                continue;
            }
            it.ei[k] = (int)(r % h.nel[k]);
            r /= h.nel[k];
        }
        out.push_back(it);
    }
}

// patch-local index of boundary-layer function jb of a side item
int64_t layer_function(const HostPatch &h, const gsx::BdryItem &it, int jb)
{
    const int dir = (it.side - 1) / 2, par = (it.side - 1) % 2, p1 = h.p[0] + 1;
    int bi[3] = {0, 0, 0}, nt = 0;
    for (int k = 0; k < h.d; ++k) {
        if (k == dir) {
            bi[k] = par ? h.p[0] : 0;
            continue;
        }
        bi[k] = nt == 0 ? jb % p1 : jb / p1;
        ++nt;
    }
    int64_t m = 0, str = 1;
    for (int k = 0; k < h.d; ++k) {
        m += (it.ei[k] + bi[k]) * str;
        str *= h.nb[k];
    }
    return m;
}

// boundary element terms of `items` (grouped by the patches' (d, p)) to host
gsx_status bdry_terms(gsx_problem *P, const std::vector<gsx::BdryItem> &items, std::vector<double> &M,
                      std::vector<double> &F, int &nbb)
{
    gsx_ctx *c = P->ctx;
    hipStream_t s = c->stream;
    const int d = P->d, p = P->hp[items[0].patch].p[0];
    nbb = d == 3 ? (p + 1) * (p + 1) : p + 1;
    const int n = (int)items.size();
    DevPtr di = dalloc(c, sizeof(gsx::BdryItem) * n), dM = dalloc(c, sizeof(double) * (size_t)n * nbb * nbb),
           dF = dalloc(c, sizeof(double) * (size_t)n * nbb);
    if (!di || !dM || !dF) return fail(c, GSX_ERR_OOM, "device allocation (boundary terms)");
    HIP_OK(c, hipMemcpyAsync(di.get(), items.data(), sizeof(gsx::BdryItem) * n, hipMemcpyHostToDevice, s));
    HIP_OK(c, gsx::launch_bdry(d, p, dp<PatchDev>(P->d_patches), dp<gsx::BdryItem>(di), n, c->gt, dp<double>(dM),
                               dp<double>(dF), s));
    M.resize((size_t)n * nbb * nbb);
    F.resize((size_t)n * nbb);
    HIP_OK(c, hipMemcpyAsync(M.data(), dM.get(), sizeof(double) * M.size(), hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipMemcpyAsync(F.data(), dF.get(), sizeof(double) * F.size(), hipMemcpyDeviceToHost, s));
    HIP_OK(c, hipStreamSynchronize(s));
    return GSX_OK;
}

}  // namespace

// Dirichlet values by interpolation (gsAssembler::computeDirichletDofsIntpl,
// src/gsAssembler/gsAssembler.hpp:314-390): side by side in condition
// order, g at the grid of the side's anchors (Greville abscissae of the
// tangential directions, the normal coordinate at the knot vector's end)
// mapped by the geometry, interpolated by the side's tensor basis; each side
// writes all of its functions, so a later side overwrites the functions it
// shares with an earlier one (homogeneous sides write zeros).
static gsx_status interpolate_sides(gsx_problem *P, const std::vector<gsx_bc> &dir, std::vector<double> &fixed)
{
    gsx_ctx *c = P->ctx;
    const int d = P->d;
    for (const gsx_bc &b : dir) {
        const HostPatch &h = P->hp[b.patch];
        const int nd = (b.side - 1) / 2, par = (b.side - 1) % 2;
        int od[2] = {0, 0}, no = 0;
        for (int k = 0; k < d; ++k)
            if (k != nd) od[no++] = k;
        const int n0 = h.nb[od[0]], n1 = d == 3 ? h.nb[od[1]] : 1;
        auto local = [&](int j0, int j1) {
            int idx[3] = {0, 0, 0};
            idx[nd] = par ? h.nb[nd] - 1 : 0;
            idx[od[0]] = j0;
            if (d == 3) idx[od[1]] = j1;
            return (int64_t)idx[0] + (int64_t)h.nb[0] * (idx[1] + (int64_t)h.nb[1] * idx[2]);
        };
        auto put = [&](int64_t l, double v) -> bool {
            const int g = P->dofs[P->boff[b.patch] + l] - P->nfree;
            if (g < 0 || g >= (int)fixed.size()) return false;
            fixed[g] = v;
            return true;
        };
        if (b.kind == GSX_BCF_ZERO) {
            for (int j1 = 0; j1 < n1; ++j1)
                for (int j0 = 0; j0 < n0; ++j0)
                    if (!put(local(j0, j1), 0.0)) return fail(c, GSX_ERR_ARG, "Dirichlet side function not eliminated");
            continue;
        }
        const std::vector<double> *U[3];
        for (int k = 0; k < d; ++k) U[k] = &P->knots_h[3 * b.patch + k];
        std::vector<double> gr[2];
        gr[0] = gsx::dense::greville(*U[od[0]], h.p[od[0]]);
        if (d == 3) gr[1] = gsx::dense::greville(*U[od[1]], h.p[od[1]]);
        const double un = par ? U[nd]->back() : U[nd]->front();
        // g at the anchor grid: V[j1][j0]
        std::vector<double> V((size_t)n0 * n1);
        const std::vector<double> &C = P->ctrl_h[b.patch];
        std::vector<double> Nv[3];
        std::vector<int> span(3);
        for (int j1 = 0; j1 < n1; ++j1)
            for (int j0 = 0; j0 < n0; ++j0) {
                double u[3] = {0, 0, 0};
                u[nd] = un;
                u[od[0]] = gr[0][j0];
                if (d == 3) u[od[1]] = gr[1][j1];
                for (int k = 0; k < d; ++k) {
                    const int p = h.p[k];
                    Nv[k].assign(p + 1, 0.0);
                    std::vector<double> dN(p + 1);
                    span[k] = gsx::dense::find_span(*U[k], p, u[k]);
                    gsx::dense::bspline_ders(*U[k], p, span[k], u[k], Nv[k].data(), dN.data());
                }
                double x[3] = {0, 0, 0};
                const int p0 = h.p[0], p1 = h.p[1], p2 = d == 3 ? h.p[2] : 0;
                for (int c2 = 0; c2 <= p2; ++c2)
                    for (int c1 = 0; c1 <= p1; ++c1)
                        for (int c0 = 0; c0 <= p0; ++c0) {
                            const double w = Nv[0][c0] * Nv[1][c1] * (d == 3 ? Nv[2][c2] : 1.0);
                            const int64_t m = (span[0] - p0 + c0) +
                                              (int64_t)h.nb[0] * ((span[1] - p1 + c1) +
                                                                  (int64_t)h.nb[1] * (d == 3 ? span[2] - p2 + c2 : 0));
                            for (int r = 0; r < d; ++r) x[r] += w * C[(size_t)m * d + r];
                        }
                double g;
                if (b.kind == GSX_BCF_MANUFACTURED) {
                    g = 1.0;
                    for (int r = 0; r < d; ++r) g *= std::sin(M_PI * x[r]);
                } else {
                    g = b.value[0];
                    for (int r = 0; r < d; ++r) g += b.value[1 + r] * x[r];
                }
                V[(size_t)j1 * n0 + j0] = g;
            }
        // collocation C_t[i][j] = N_j(anchor_i); coefs = C_1^-1 V C_0^-T
        auto colloc = [&](int t, std::vector<double> &M) {
            const int k = od[t], p = h.p[k], n = h.nb[k];
            M.assign((size_t)n * n, 0.0);
            std::vector<double> N(p + 1), dN(p + 1);
            for (int i = 0; i < n; ++i) {
                const int sp = gsx::dense::find_span(*U[k], p, gr[t][i]);
                gsx::dense::bspline_ders(*U[k], p, sp, gr[t][i], N.data(), dN.data());
                for (int a = 0; a <= p; ++a) M[(size_t)i * n + sp - p + a] = N[a];
            }
        };
        std::vector<double> M0, W((size_t)n0 * n1);
        colloc(0, M0);
        for (int j1 = 0; j1 < n1; ++j1)
            for (int j0 = 0; j0 < n0; ++j0) W[(size_t)j0 * n1 + j1] = V[(size_t)j1 * n0 + j0];
        if (!gsx::dense::lu_solve(n0, M0, W, n1)) return fail(c, GSX_ERR_ARG, "singular boundary collocation");
        // W[j0][j1] = (C_0^-1 V^T)[j0][j1]
        if (d == 3) {
            std::vector<double> M1, Z((size_t)n1 * n0);
            colloc(1, M1);
            for (int j0 = 0; j0 < n0; ++j0)
                for (int j1 = 0; j1 < n1; ++j1) Z[(size_t)j1 * n0 + j0] = W[(size_t)j0 * n1 + j1];
            if (!gsx::dense::lu_solve(n1, M1, Z, n0)) return fail(c, GSX_ERR_ARG, "singular boundary collocation");
            for (int j1 = 0; j1 < n1; ++j1)
                for (int j0 = 0; j0 < n0; ++j0) W[(size_t)j0 * n1 + j1] = Z[(size_t)j1 * n0 + j0];
        }
        for (int j1 = 0; j1 < n1; ++j1)
            for (int j0 = 0; j0 < n0; ++j0)
                if (!put(local(j0, j1), W[(size_t)j0 * n1 + j1]))
                    return fail(c, GSX_ERR_ARG, "Dirichlet side function not eliminated");
    }
    return GSX_OK;
}

extern "C" gsx_status gsx_problem_set_bc(gsx_problem *P, int32_t nbc, const gsx_bc *bc, int32_t dir_values)
{
    if (!P || nbc < 0 || (nbc > 0 && !bc)) return GSX_ERR_ARG;
    gsx_ctx *c = P->ctx;
    if (P->geo) return fail(c, GSX_ERR_ARG, "boundary data need the problem's own geometry");
    hipStream_t s = c->stream;
    if (dir_values != GSX_DIR_HOMOGENEOUS && dir_values != GSX_DIR_L2PROJECTION &&
        dir_values != GSX_DIR_L2PROJECTION_INHOM && dir_values != GSX_DIR_INTERPOLATION)
        return fail(c, GSX_ERR_ARG, "unknown Dirichlet value method");
    const int d = P->d;
    std::vector<std::pair<int, int>> dset, iset;
    for (size_t k = 0; k + 1 < P->dir_sides_h.size(); k += 2) dset.push_back({P->dir_sides_h[k], P->dir_sides_h[k + 1]});
    for (size_t k = 0; k + 7 < P->ifaces_h.size(); k += 8) {
        iset.push_back({P->ifaces_h[k], P->ifaces_h[k + 1]});
        iset.push_back({P->ifaces_h[k + 2], P->ifaces_h[k + 3]});
    }
    auto has = [](const std::vector<std::pair<int, int>> &v, int a, int b) {
        return std::find(v.begin(), v.end(), std::make_pair(a, b)) != v.end();
    };
    // Dirichlet data per eliminated side (problem order), Neumann records
    std::vector<gsx_bc> dir(dset.size()), neu;
    for (size_t k = 0; k < dset.size(); ++k) {
        dir[k] = gsx_bc{};
        dir[k].patch = dset[k].first;
        dir[k].side = dset[k].second;
        dir[k].type = GSX_BC_DIRICHLET;
        dir[k].kind = GSX_BCF_ZERO;
    }
    std::vector<std::pair<int, int>> seen;
    for (int k = 0; k < nbc; ++k) {
        const gsx_bc &b = bc[k];
        if (b.patch < 0 || b.patch >= P->npatch || b.side < 1 || b.side > 2 * d ||
            (b.type != GSX_BC_DIRICHLET && b.type != GSX_BC_NEUMANN) || b.kind < GSX_BCF_ZERO ||
            b.kind > GSX_BCF_AFFINE)
            return fail(c, GSX_ERR_ARG, "invalid boundary condition record");
        if (has(seen, b.patch, b.side)) return fail(c, GSX_ERR_ARG, "two boundary conditions on one side");
        seen.push_back({b.patch, b.side});
        if (b.type == GSX_BC_DIRICHLET) {
            if (!has(dset, b.patch, b.side))
                return fail(c, GSX_ERR_ARG, "Dirichlet data on a side that the problem does not eliminate");
            for (size_t j = 0; j < dset.size(); ++j)
                if (dset[j] == std::make_pair(b.patch, b.side)) dir[j] = b;
        } else {
            if (has(dset, b.patch, b.side) || has(iset, b.patch, b.side))
                return fail(c, GSX_ERR_ARG, "Neumann data on an eliminated or interface side");
            neu.push_back(b);
        }
    }
    for (const gsx_bc &b : dir)
        if (P->hp[b.patch].p[0] > gsx::kMaxDeg) return fail(c, GSX_ERR_ARG, "degree not supported");
    P->d_fixed.reset();
    P->d_neu.reset();
    const int64_t nfix = P->nglued - P->nfree;
    bool inhom = false;
    for (const gsx_bc &b : dir) inhom = inhom || b.kind != GSX_BCF_ZERO;
    if (dir_values == GSX_DIR_INTERPOLATION && inhom && nfix > 0) {
        std::vector<double> fixed((size_t)nfix, 0.0);
        GSX_TRY(interpolate_sides(P, dir, fixed));
        P->d_fixed = dalloc(c, sizeof(double) * nfix);
        if (!P->d_fixed) return fail(c, GSX_ERR_OOM, "device allocation (Dirichlet values)");
        HIP_OK(c, hipMemcpyAsync(P->d_fixed.get(), fixed.data(), sizeof(double) * nfix, hipMemcpyHostToDevice, s));
        HIP_OK(c, hipStreamSynchronize(s));
    } else if (dir_values != GSX_DIR_HOMOGENEOUS && inhom && nfix > 0) {
        // boundary mass matrix and load over every Dirichlet side
        // (gsDirichletValuesByL2Projection, src/gsAssembler/gsDirichletValues.h:257-435)
        std::vector<std::map<int, double>> rows((size_t)nfix);
        std::vector<double> rhs((size_t)nfix, 0.0);
        for (const Group &g : P->groups) {
            std::vector<gsx::BdryItem> items;
            for (const gsx_bc &b : dir)
                if (std::find(g.patches.begin(), g.patches.end(), b.patch) != g.patches.end() &&
                    !(dir_values == GSX_DIR_L2PROJECTION_INHOM && b.kind == GSX_BCF_ZERO))
                    side_items(P->hp[b.patch], b.patch, b, items);
            if (items.empty()) continue;
            std::vector<double> M, F;
            int nbb = 0;
            GSX_TRY(bdry_terms(P, items, M, F, nbb));
            for (size_t t = 0; t < items.size(); ++t) {
                const HostPatch &h = P->hp[items[t].patch];
                std::vector<int> gl(nbb);
                for (int j = 0; j < nbb; ++j) {
                    gl[j] = P->dofs[P->boff[items[t].patch] + layer_function(h, items[t], j)] - P->nfree;
                    if (gl[j] < 0 || gl[j] >= nfix) return fail(c, GSX_ERR_ARG, "Dirichlet side function not eliminated");
                }
                for (int i = 0; i < nbb; ++i) {
                    for (int j = 0; j < nbb; ++j) rows[gl[i]][gl[j]] += M[(t * nbb + i) * nbb + j];
                    rhs[gl[i]] += F[t * nbb + i];
                }
            }
        }
        std::vector<int> rp((size_t)nfix + 1, 0), ci;
        std::vector<double> va;
        // functions on no projected side (homogeneous sides skipped): an
        // identity row keeps their value 0
        for (int64_t i = 0; i < nfix; ++i)
            if (rows[i].empty()) rows[i][(int)i] = 1.0;
        for (int64_t i = 0; i < nfix; ++i) {
            for (const auto &kv : rows[i]) {
                ci.push_back(kv.first);
                va.push_back(kv.second);
            }
            rp[i + 1] = (int)ci.size();
        }
        gsx_matrix *Mb = nullptr;
        GSX_TRY(gsx_matrix_from_csr(c, (int32_t)nfix, rp.data(), ci.data(), va.data(), 0, &Mb));
        std::unique_ptr<gsx_matrix, gsx_status (*)(gsx_matrix *)> hold(Mb, gsx_matrix_destroy);
        DevPtr db = dalloc(c, sizeof(double) * nfix), dx = dalloc(c, sizeof(double) * nfix);
        if (!db || !dx) return fail(c, GSX_ERR_OOM, "device allocation (Dirichlet values)");
        HIP_OK(c, hipMemcpyAsync(db.get(), rhs.data(), sizeof(double) * nfix, hipMemcpyHostToDevice, s));
        HIP_OK(c, gsx::launch_fill_f64(dp<double>(dx), nfix, 0.0, s));
        // gsSparseSolver::CGDiagonal with Eigen's default tolerance (machine
        // epsilon): Jacobi-CG to rounding level (a boundary mass matrix is
        // well conditioned: a few dozen iterations)
        int32_t it = 0;
        double rel = 0.0;
        const gsx_status e = gsx_cg_solve(Mb, GSX_PREC_JACOBI, 1.0, nullptr, GSX_FP64, dp<double>(db), dp<double>(dx),
                                          1e-14, 2000, &it, &rel);
        if (e != GSX_OK && !(e == GSX_ERR_NOT_CONVERGED && rel < 1e-12)) return e;
        HIP_OK(c, hipStreamSynchronize(s));
        P->d_fixed = std::move(dx);
    }
    bool any_neu = false;
    for (const gsx_bc &b : neu) any_neu = any_neu || b.kind != GSX_BCF_ZERO;
    if (any_neu && P->nfree > 0) {
        std::vector<double> load((size_t)P->nfree, 0.0);
        // records in order, side elements in iterator order
        for (const gsx_bc &b : neu) {
            std::vector<gsx::BdryItem> items;
            side_items(P->hp[b.patch], b.patch, b, items);
            std::vector<double> M, F;
            int nbb = 0;
            GSX_TRY(bdry_terms(P, items, M, F, nbb));
            for (size_t t = 0; t < items.size(); ++t) {
                const HostPatch &h = P->hp[items[t].patch];
                for (int j = 0; j < nbb; ++j) {
                    const int g = P->dofs[P->boff[items[t].patch] + layer_function(h, items[t], j)];
                    if (g < P->nfree) load[g] += F[t * nbb + j];
                }
            }
        }
        P->d_neu = dalloc(c, sizeof(double) * P->nfree);
        if (!P->d_neu) return fail(c, GSX_ERR_OOM, "device allocation (Neumann load)");
        HIP_OK(c, hipMemcpyAsync(P->d_neu.get(), load.data(), sizeof(double) * P->nfree, hipMemcpyHostToDevice, s));
        HIP_OK(c, hipStreamSynchronize(s));
    }
    return GSX_OK;
}

extern "C" gsx_status gsx_problem_fixed_values(const gsx_problem *P, double *fixed, int32_t count)
{
    if (!P || (count > 0 && !fixed)) return GSX_ERR_ARG;
    const int64_t nfix = P->nglued - P->nfree;
    if (count != nfix) return GSX_ERR_ARG;
    if (!P->d_fixed) {
        std::fill(fixed, fixed + count, 0.0);
        return GSX_OK;
    }
    gsx_ctx *c = P->ctx;
    HIP_OK(c, hipMemcpyAsync(fixed, P->d_fixed.get(), sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return GSX_OK;
}

extern "C" gsx_status gsx_assemble_bdr(gsx_problem *P, double *rhs)
{
    if (!P || !rhs) return GSX_ERR_ARG;
    if (!P->d_neu) return GSX_OK;
    gsx_ctx *c = P->ctx;
    HIP_OK(c, gsx::launch_axpby(P->nfree, 1.0, dp<double>(P->d_neu), 1.0, rhs, c->stream));
    return GSX_OK;
}
