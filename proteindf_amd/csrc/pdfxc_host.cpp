// Host side of the pdfxc engine: the C ABI, the AO shell table, grid-block
// partitioning with per-block AO screening, per-class work lists, device
// buffers and grid generation.  Kernels live in pdfxc_kernels.hip.
#include <algorithm>
#include <functional>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "pdfxc_internal.hpp"
#include "xc_functional.hpp"

namespace pdfxc {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return PDFXC_ERR_HIP;
}

static int arg_error(const std::string& msg) {
    set_error(msg);
    return PDFXC_ERR_ARGUMENT;
}

template <typename T>
static int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                  " bytes: " + hipGetErrorString(e));
        return PDFXC_ERR_MEMORY;
    }
    return PDFXC_OK;
}

template <typename T>
static void dfree(T*& p) {
    if (p) (void)hipFree(reinterpret_cast<void*>(p));
    p = nullptr;
}

// ---------------------------------------------------------------------------
// Contracted-GTO normalisation, as TlOrbitalInfoObject::setCGTO forms the
// coefficients (TlOrbitalInfoObject.cpp:38-106): c_k = d_k N(a_k) N_c with the
// primitive norm of Fl_Gto::getNormalized (Fl_Gto.cpp:644-708) and the
// contraction norm of getNormalizedfactor (:558-641), for the angular
// exponents (l,m,n) = (0,0,0), (1,0,0), (1,1,0) of s, p, d (the five d
// functions share the dxy norm).  Primitives end sorted by exponent.
// ---------------------------------------------------------------------------
static double odd_factorial2(int x) {  // (x)!! for odd x >= -1
    double a = 1.0;
    for (; x > 1; x -= 2) a *= x;
    return a;
}

static void lmn_of_shell(int shell, int* l, int* m, int* n) {
    *l = *m = *n = 0;
    if (shell >= 1) *l = 1;
    if (shell >= 2) *m = 1;
}

static double primitive_norm(double a, int l, int m, int n) {
    const int L = l + m + n;
    return std::pow(2.0, (double)L) *
           std::pow(odd_factorial2(2 * l - 1) * odd_factorial2(2 * m - 1) * odd_factorial2(2 * n - 1),
                    -0.5) *
           std::pow(2.0 / kPi, 0.75) * std::pow(a, (L + 1.5) / 2.0);
}

}  // namespace pdfxc

using namespace pdfxc;

extern "C" {

void pdfxc_default_opts(pdfxc_opts* o) {
    if (!o) return;
    o->density_cutoff = 1.0e-16;
    o->ao_screen = 0.0;
    o->precision = PDFXC_FP64;
    o->block_points = kBlockPts;
    o->scratch_bytes = 0;
}

const char* pdfxc_last_error(void) { return g_last_error.c_str(); }

int pdfxc_normalize_cgto(int shell, int nprim, const double* exps, const double* coefs,
                         double* out_exps, double* out_coefs) {
    if (shell < 0 || shell > 2) {
        set_error("pdfxc_normalize_cgto: shell must be 0 (s), 1 (p) or 2 (d)");
        return PDFXC_ERR_UNSUPPORTED;
    }
    if (nprim <= 0 || nprim > kMaxPrim || !exps || !coefs || !out_exps || !out_coefs)
        return arg_error("pdfxc_normalize_cgto: bad arguments");
    int l, m, n;
    lmn_of_shell(shell, &l, &m, &n);
    const double pw = (double)(l + m + n) + 1.5;
    double s = 0.0;
    for (int a = 0; a < nprim; ++a)
        for (int b = 0; b < nprim; ++b)
            s += coefs[a] * coefs[b] * primitive_norm(exps[a], l, m, n) *
                 primitive_norm(exps[b], l, m, n) * std::pow(exps[a] + exps[b], -pw);
    s *= odd_factorial2(2 * l - 1) * odd_factorial2(2 * m - 1) * odd_factorial2(2 * n - 1);
    s *= std::pow(2.0, -(double)(l + m + n));
    s *= std::pow(kPi, 1.5);
    const double nc = std::sqrt(1.0 / s);
    std::vector<int> order(nprim);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return exps[x] < exps[y]; });
    for (int k = 0; k < nprim; ++k) {
        const int o = order[k];
        out_exps[k] = exps[o];
        out_coefs[k] = coefs[o] * primitive_norm(exps[o], l, m, n) * nc;
    }
    return PDFXC_OK;
}

int pdfxc_ctx_create(pdfxc_ctx** out, int n_aos, const double* ao_center, const int* ao_type,
                     const int* ao_nprim, const double* ao_exps, const double* ao_coefs,
                     int max_prim, const pdfxc_opts* opts) {
    if (!out) return arg_error("pdfxc_ctx_create: null output");
    *out = nullptr;
    if (n_aos <= 0 || !ao_center || !ao_type || !ao_nprim || !ao_exps || !ao_coefs || max_prim <= 0)
        return arg_error("pdfxc_ctx_create: bad arguments");
    Context* c = new Context();
    pdfxc_default_opts(&c->opts);
    if (opts) c->opts = *opts;
    if (c->opts.block_points != kBlockPts) {
        delete c;
        set_error("pdfxc_ctx_create: only 128 points per block are supported");
        return PDFXC_ERR_UNSUPPORTED;
    }
    if (c->opts.precision != PDFXC_FP64 && c->opts.precision != PDFXC_SPLIT_BF16 &&
        c->opts.precision != PDFXC_SPLIT_BF16X3 && c->opts.precision != PDFXC_SPLIT_K_BF16X3 &&
        c->opts.precision != PDFXC_SPLIT_BFP && c->opts.precision != PDFXC_SPLIT_K_F16) {
        delete c;
        set_error("pdfxc_ctx_create: unknown precision mode " + std::to_string(c->opts.precision));
        return PDFXC_ERR_ARGUMENT;
    }
    c->n_aos = n_aos;
    // group AOs into shells: type 0 | 1,2,3 | 4..8 on one centre, same primitives
    int i = 0;
    while (i < n_aos) {
        const int t = ao_type[i];
        int l, ncomp;
        if (t == 0) {
            l = 0;
            ncomp = 1;
        } else if (t == 1) {
            l = 1;
            ncomp = 3;
        } else if (t == 4) {
            l = 2;
            ncomp = 5;
        } else {
            delete c;
            set_error("pdfxc_ctx_create: AO " + std::to_string(i) + " of type " + std::to_string(t) +
                      " does not start a shell (s=0, p=1..3, d=4..8 consecutive)");
            return (t < 0 || t > 8) ? PDFXC_ERR_UNSUPPORTED : PDFXC_ERR_ARGUMENT;
        }
        if (i + ncomp > n_aos) {
            delete c;
            return arg_error("pdfxc_ctx_create: truncated shell at the end of the AO table");
        }
        Shell s;
        s.first_ao = i;
        s.l = l;
        s.nprim = ao_nprim[i];
        if (s.nprim <= 0 || s.nprim > kMaxPrim || s.nprim > max_prim) {
            delete c;
            return arg_error("pdfxc_ctx_create: bad primitive count at AO " + std::to_string(i));
        }
        for (int d = 0; d < 3; ++d) s.center[d] = ao_center[3 * i + d];
        for (int k = 0; k < s.nprim; ++k) {
            s.exps[k] = ao_exps[(size_t)i * max_prim + k];
            s.coefs[k] = ao_coefs[(size_t)i * max_prim + k];
            if (!(s.exps[k] > 0.0)) {
                delete c;
                return arg_error("pdfxc_ctx_create: non-positive exponent at AO " + std::to_string(i));
            }
        }
        for (int m = 1; m < ncomp; ++m) {
            const int j = i + m;
            bool ok = ao_type[j] == t + m && ao_nprim[j] == s.nprim;
            for (int d = 0; d < 3 && ok; ++d) ok = ao_center[3 * j + d] == s.center[d];
            for (int k = 0; k < s.nprim && ok; ++k)
                ok = ao_exps[(size_t)j * max_prim + k] == s.exps[k] &&
                     ao_coefs[(size_t)j * max_prim + k] == s.coefs[k];
            if (!ok) {
                delete c;
                return arg_error("pdfxc_ctx_create: AO " + std::to_string(j) +
                                 " does not continue the shell starting at AO " + std::to_string(i));
            }
        }
        // primitives in ascending exponent order, as setCGTO keeps them
        // (TlOrbitalInfoObject.cpp:38-106): the kernels stop at the first
        // primitive past the alpha r^2 > 30 cut
        {
            int ord[kMaxPrim];
            for (int k = 0; k < s.nprim; ++k) ord[k] = k;
            std::stable_sort(ord, ord + s.nprim, [&](int x, int y) { return s.exps[x] < s.exps[y]; });
            double e2[kMaxPrim], c2[kMaxPrim];
            for (int k = 0; k < s.nprim; ++k) {
                e2[k] = s.exps[ord[k]];
                c2[k] = s.coefs[ord[k]];
            }
            for (int k = 0; k < s.nprim; ++k) {
                s.exps[k] = e2[k];
                s.coefs[k] = c2[k];
            }
        }
        c->shells.push_back(s);
        i += ncomp;
    }
    {
        const int rc0 = configure_kernels();
        if (rc0 != PDFXC_OK) {
            delete c;
            return rc0;
        }
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int k = 0; k < 6 && e == hipSuccess; ++k) e = hipEventCreate(&c->ev[k]);
    if (e != hipSuccess) {
        const int rc = hip_fail(e, "pdfxc_ctx_create: stream/event creation");
        delete c;
        return rc;
    }
    std::vector<DeviceShell> ds(c->shells.size());
    for (size_t k = 0; k < c->shells.size(); ++k) {
        const Shell& s = c->shells[k];
        DeviceShell& d = ds[k];
        std::memset(&d, 0, sizeof(d));
        d.cx = s.center[0];
        d.cy = s.center[1];
        d.cz = s.center[2];
        d.l = s.l;
        d.nprim = s.nprim;
        d.first_ao = s.first_ao;
        for (int q = 0; q < s.nprim; ++q) {
            d.exps[q] = s.exps[q];
            d.coefs[q] = s.coefs[q];
        }
    }
    int maxp = 1;
    for (const Shell& sh : c->shells) maxp = std::max(maxp, sh.nprim);
    c->rec_stride = 4 + 2 * maxp;
    std::vector<double> rec(c->shells.size() * (size_t)c->rec_stride, 0.0);
    for (size_t k = 0; k < c->shells.size(); ++k) {
        const Shell& sh = c->shells[k];
        double* r = rec.data() + k * c->rec_stride;
        r[0] = sh.center[0];
        r[1] = sh.center[1];
        r[2] = sh.center[2];
        r[3] = (double)(sh.l + 4 * sh.nprim);
        for (int q = 0; q < sh.nprim; ++q) {
            r[4 + 2 * q] = sh.exps[q];
            r[5 + 2 * q] = sh.coefs[q];
        }
    }
    int rc = dalloc(&c->d_shells, ds.size());
    if (rc == PDFXC_OK) rc = dalloc(&c->d_shrec, rec.size());
    if (rc == PDFXC_OK) {
        e = hipMemcpy(c->d_shrec, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_ctx_create: shell records upload");
    }
    if (rc == PDFXC_OK) rc = dalloc(&c->d_P, (size_t)n_aos * n_aos);
    if (rc == PDFXC_OK) rc = dalloc(&c->d_K, (size_t)n_aos * n_aos);
    if (rc == PDFXC_OK) rc = dalloc(&c->d_E, 1);
    if (rc == PDFXC_OK) rc = dalloc(&c->d_queue, kMaxQueues);
    if (rc == PDFXC_OK) {
        e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_ctx_create: compute unit count");
    }
    if (rc == PDFXC_OK && c->opts.precision != PDFXC_FP64) {
        const int S = prec_slices(c->opts.precision);
        c->split_nc = (n_aos + 7) / 8 * 8;
        // split AO order: atoms (runs of shells on one centre) in input order,
        // each atom's shells by most diffuse exponent first, so that the shells
        // a block keeps of a distant atom form one run of columns
        {
            std::vector<int> order(c->shells.size());
            std::iota(order.begin(), order.end(), 0);
            size_t a0 = 0;
            while (a0 < c->shells.size()) {
                size_t a1 = a0 + 1;
                while (a1 < c->shells.size() &&
                       std::equal(c->shells[a1].center, c->shells[a1].center + 3, c->shells[a0].center))
                    ++a1;
                std::stable_sort(order.begin() + a0, order.begin() + a1,
                                 [&](int x, int y) { return c->shells[x].exps[0] < c->shells[y].exps[0]; });
                a0 = a1;
            }
            c->split_perm.clear();
            for (int k : order)
                for (int m = 0; m < 2 * c->shells[k].l + 1; ++m) c->split_perm.push_back(c->shells[k].first_ao + m);
        }
        (void)S;
        rc = dalloc(&c->d_Ps, (size_t)psplit_elems(c->opts.precision, n_aos));
        if (rc == PDFXC_OK) rc = dalloc(&c->d_perm, (size_t)n_aos);
        if (rc == PDFXC_OK) {
            e = hipMemcpy(c->d_perm, c->split_perm.data(), sizeof(int) * n_aos, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = hip_fail(e, "pdfxc_ctx_create: split order upload");
        }
    }
    if (rc == PDFXC_OK) {
        e = hipMemcpy(c->d_shells, ds.data(), ds.size() * sizeof(DeviceShell), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_ctx_create: shell upload");
    }
    if (rc == PDFXC_OK) {
        e = hipHostMalloc(reinterpret_cast<void**>(&c->h_pinned),
                          sizeof(double) * (size_t)n_aos * n_aos, hipHostMallocDefault);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_ctx_create: pinned staging");
    }
    if (rc != PDFXC_OK) {
        pdfxc_ctx_destroy(reinterpret_cast<pdfxc_ctx*>(c));
        return rc;
    }
    *out = reinterpret_cast<pdfxc_ctx*>(c);
    return PDFXC_OK;
}

static void free_grid_buffers(Context* c) {
    dfree(c->d_pts);
    dfree(c->d_blocks);
    dfree(c->d_block_shells);
    dfree(c->d_block_cols);
    dfree(c->d_block_shell_col);
    dfree(c->d_block_sorder);
    dfree(c->d_rho_work);
    dfree(c->d_k_work);
    dfree(c->d_coef);
    dfree(c->d_rho);
    dfree(c->d_coef_b);
    dfree(c->d_rho_b);
    dfree(c->d_eblock);
    dfree(c->d_sblocks);
    dfree(c->d_chunks);
    dfree(c->d_sshell_col);
    dfree(c->d_srho_work);
    dfree(c->d_kg_work);
    dfree(c->d_kgroups);
    dfree(c->d_kg_blocks);
    dfree(c->d_kg_shells);
    dfree(c->d_kg_shell_col);
    dfree(c->d_kg_sorder);
    dfree(c->d_kg_cols);
    dfree(c->d_kg_masks);
}

int pdfxc_ctx_destroy(pdfxc_ctx* ctx) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return PDFXC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_grid_buffers(c);
    dfree(c->d_shells);
    dfree(c->d_shrec);
    dfree(c->d_P);
    dfree(c->d_K);
    dfree(c->d_queue);
    dfree(c->d_P2);
    dfree(c->d_K2);
    dfree(c->d_E);
    dfree(c->d_Ps);
    dfree(c->d_perm);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PDFXC_OK;
}

void* pdfxc_stream(pdfxc_ctx* ctx) {
    Context* c = reinterpret_cast<Context*>(ctx);
    return c ? reinterpret_cast<void*>(c->stream) : nullptr;
}

}  // extern "C"

namespace pdfxc {

// ---------------------------------------------------------------------------
// Grid blocks: recursive bisection of the point set along its widest extent,
// cutting at a multiple of 128 so that every block but the last is full.
// ---------------------------------------------------------------------------
static void partition_points(const Context& c, std::vector<int64_t>& order,
                             std::vector<std::pair<int64_t, int64_t>>& leaves) {
    order.resize(c.n_pts);
    std::iota(order.begin(), order.end(), 0);
    std::vector<std::pair<int64_t, int64_t>> stack;
    stack.push_back({0, c.n_pts});
    const double* X[3] = {c.gx.data(), c.gy.data(), c.gz.data()};
    while (!stack.empty()) {
        const auto [lo, hi] = stack.back();
        stack.pop_back();
        const int64_t n = hi - lo;
        if (n <= kBlockPts) {
            leaves.push_back({lo, hi});
            continue;
        }
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int64_t k = lo; k < hi; ++k)
            for (int d = 0; d < 3; ++d) {
                const double v = X[d][order[k]];
                mn[d] = std::min(mn[d], v);
                mx[d] = std::max(mx[d], v);
            }
        int dim = 0;
        for (int d = 1; d < 3; ++d)
            if (mx[d] - mn[d] > mx[dim] - mn[dim]) dim = d;
        const int64_t nch = (n + kBlockPts - 1) / kBlockPts;
        const int64_t mid = lo + (nch / 2) * kBlockPts;
        const double* v = X[dim];
        std::nth_element(order.begin() + lo, order.begin() + mid, order.begin() + hi,
                         [v](int64_t a, int64_t b) { return v[a] < v[b] || (v[a] == v[b] && a < b); });
        // right first so that leaves come out left to right
        stack.push_back({mid, hi});
        stack.push_back({lo, mid});
    }
}

// Upper bound of |phi| and |grad phi| of a shell over r >= d: sum over
// primitives of |c| (r^l + l r^(l-1) + 2a r^(l+1)) e^(-a r^2), each term at
// its own maximiser max(d, sqrt(m / 2a)).
static double shell_bound(const Shell& s, double d) {
    double b = 0.0;
    for (int k = 0; k < s.nprim; ++k) {
        const double a = s.exps[k];
        auto term = [&](int m) {
            if (m < 0) return 0.0;
            const double r = std::max(d, std::sqrt(m / (2.0 * a)));
            return std::pow(r, (double)m) * std::exp(-a * r * r);
        };
        b += std::fabs(s.coefs[k]) * (term(s.l) + 2.0 * s.l * term(s.l - 1) + 2.0 * a * term(s.l + 1));
    }
    return b;
}

// Drop every grid-dependent host list and device buffer: a context left like
// this refuses builds with PDFXC_ERR_NO_GRID.
static void reset_grid(Context& c) {
    free_grid_buffers(&c);
    c.n_pts = 0;
    c.n_slots = 0;
    c.gx.clear();
    c.gy.clear();
    c.gz.clear();
    c.gw.clear();
    c.gen_atom.clear();
    c.blocks.clear();
    c.block_shells.clear();
    c.block_cols.clear();
    c.block_shell_col.clear();
    c.block_sorder.clear();
    c.slot_of_point.clear();
    c.rho_launches.clear();
    c.kmat_launches.clear();
    c.sblocks.clear();
    c.chunks.clear();
    c.sshell_col.clear();
    c.srho_launches.clear();
    c.sk_launches.clear();
    c.k_fp64_launches.clear();
    c.h_rho_work.clear();
    c.h_k_work.clear();
    c.h_srho_work.clear();
    c.h_kg_work.clear();
    c.kgroups.clear();
    c.kg_blocks.clear();
    c.kg_shells.clear();
    c.kg_shell_col.clear();
    c.kg_sorder.clear();
    c.kg_cols.clear();
    c.kg_masks.clear();
    c.flops = 0.0;
    c.mean_nao = 0.0;
    c.max_nao = 0;
    c.n_k_fp64_blocks = 0;
}

static int build_blocks_impl(Context& c);

// Blocks, work lists and device buffers of the context's grid; on failure the
// context is left with no grid (nothing half-built stays reachable).
static int build_blocks(Context& c) {
    const int rc = build_blocks_impl(c);
    if (rc != PDFXC_OK) {
        const std::string msg = g_last_error;
        reset_grid(c);
        set_error(msg);
    }
    return rc;
}

static int build_blocks_impl(Context& c) {
    std::vector<int64_t> order;
    std::vector<std::pair<int64_t, int64_t>> leaves;
    partition_points(c, order, leaves);
    const int nb = (int)leaves.size();
    const int nsh = (int)c.shells.size();
    struct Tmp {
        int64_t lo, hi;
        std::vector<int> sh;
        int nao;
        double cost;
    };
    std::vector<Tmp> tmp(nb);
    const double screen = c.opts.ao_screen;
    for (int b = 0; b < nb; ++b) {
        Tmp& t = tmp[b];
        t.lo = leaves[b].first;
        t.hi = leaves[b].second;
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int64_t k = t.lo; k < t.hi; ++k) {
            const int64_t p = order[k];
            const double v[3] = {c.gx[p], c.gy[p], c.gz[p]};
            for (int d = 0; d < 3; ++d) {
                mn[d] = std::min(mn[d], v[d]);
                mx[d] = std::max(mx[d], v[d]);
            }
        }
        t.nao = 0;
        for (int s = 0; s < nsh; ++s) {
            const Shell& S = c.shells[s];
            // distance from the shell's centre to the block's bounding box: no
            // point of the block is nearer (shrunk by a relative 1e-12 for rounding)
            double d2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double e = std::max(0.0, std::max(mn[d] - S.center[d], S.center[d] - mx[d]));
                d2 += e * e;
            }
            const double dmin = std::sqrt(d2) * (1.0 - 1e-12);
            // the reference skips a primitive with a r^2 > 30: a shell whose most
            // diffuse primitive is cut at the nearest point is exactly zero here
            if (S.exps[0] * dmin * dmin > 30.0) continue;
            if (screen > 0.0 && shell_bound(S, dmin) < screen) continue;
            t.sh.push_back(s);
            t.nao += 2 * S.l + 1;
        }
        const int ldn = std::max(kColPad, (t.nao + kColPad - 1) / kColPad * kColPad);
        t.cost = (double)(t.hi - t.lo) * ldn * ldn + 1.0;
    }
    // launch order: descending cost (longest blocks first)
    std::vector<int> lorder(nb);
    std::iota(lorder.begin(), lorder.end(), 0);
    std::stable_sort(lorder.begin(), lorder.end(),
                     [&](int a, int b) { return tmp[a].cost > tmp[b].cost; });
    c.blocks.assign(nb, BlockDesc());
    c.block_shells.clear();
    c.block_cols.clear();
    c.block_shell_col.clear();
    c.block_sorder.clear();
    c.n_slots = (int64_t)nb * kBlockPts;
    c.slot_of_point.assign(c.n_pts, -1);
    std::vector<double> pts(4 * (size_t)c.n_slots, 0.0);
    c.flops = 0.0;
    double sum_nao = 0.0;
    c.max_nao = 0;
    for (int li = 0; li < nb; ++li) {
        const Tmp& t = tmp[lorder[li]];
        BlockDesc& B = c.blocks[li];
        B.pt_begin = li * kBlockPts;
        B.npts = (int)(t.hi - t.lo);
        B.shell_begin = (int)c.block_shells.size();
        B.nshell = (int)t.sh.size();
        B.col_begin = (int)c.block_cols.size();
        B.nao = t.nao;
        B.ldn = std::max(kColPad, (t.nao + kColPad - 1) / kColPad * kColPad);
        B.ldp = c.n_aos;  // one molecule: P and K start at 0
        B.poff = 0;
        B.psoff = 0;
        if (B.ldn > 1024) {
            set_error("set_grid: a block keeps " + std::to_string(B.nao) +
                      " AOs; more than 1024 per block is not supported (raise ao_screen)");
            return PDFXC_ERR_UNSUPPORTED;
        }
        int col = 0;
        for (int s : t.sh) {
            c.block_shells.push_back(s);
            c.block_shell_col.push_back(col);
            const Shell& S = c.shells[s];
            for (int m = 0; m < 2 * S.l + 1; ++m) c.block_cols.push_back(S.first_ao + m);
            col += 2 * S.l + 1;
        }
        // evaluation order: shells of equal (l, nprim) side by side, so that a
        // wave's lanes mostly run the same code path
        {
            std::vector<int> so(B.nshell);
            std::iota(so.begin(), so.end(), 0);
            std::stable_sort(so.begin(), so.end(), [&](int a, int b) {
                const Shell& A = c.shells[t.sh[a]];
                const Shell& Bs = c.shells[t.sh[b]];
                return A.l != Bs.l ? A.l < Bs.l : A.nprim < Bs.nprim;
            });
            c.block_sorder.insert(c.block_sorder.end(), so.begin(), so.end());
        }
        for (int k = 0; k < kBlockPts; ++k) {
            const int64_t slot = (int64_t)li * kBlockPts + k;
            const int64_t p = order[t.lo + std::min<int64_t>(k, B.npts - 1)];
            pts[slot] = c.gx[p];
            pts[c.n_slots + slot] = c.gy[p];
            pts[2 * c.n_slots + slot] = c.gz[p];
            pts[3 * c.n_slots + slot] = (k < B.npts) ? c.gw[p] : 0.0;
            if (k < B.npts) c.slot_of_point[p] = slot;
        }
        c.flops += 4.0 * B.npts * (double)B.nao * B.nao;
        sum_nao += B.nao;
        c.max_nao = std::max(c.max_nao, B.nao);
    }
    c.mean_nao = nb ? sum_nao / nb : 0.0;

    // k_rho work lists by template class: ldn <= 128 (64 points per
    // workgroup), <= 256 (64 when the plane fits in LDS: class 4, else 32),
    // <= 512, <= 1024 (16); the tuning may move every block up by rho_shift
    // classes (fewer points per workgroup; ldn <= 256 then takes 16)
    std::vector<int> rho_work;
    c.rho_launches.clear();
    const int RS = c.rec_stride;
    auto rho_lds = [&](const BlockDesc& B, size_t pb) {
        return sizeof(double) * (std::max<size_t>((size_t)B.ldn * pb, 4 * 512) + (size_t)B.nshell * RS) +
               sizeof(int) * (size_t)(B.ldn + 2 * B.nshell);
    };
    size_t rho64_lds_max = 0;
    if (int rc = rho64_max_dyn_lds(&rho64_lds_max)) return rc;
    auto rho_class = [&](const BlockDesc& B) {
        int base = 0;
        while ((128 << base) < B.ldn) ++base;
        const int k = std::min(3, base + c.tune.rho_shift);
        return (k == 1 && c.tune.rho_shift == 0 && rho_lds(B, 64) <= rho64_lds_max) ? 4 : k;
    };
    for (int cls = 0; cls < 5; ++cls) {
        const size_t pb = (size_t)rho_points_per_workgroup(cls);
        LaunchRange r{cls, (int)rho_work.size(), 0, 0};
        size_t maxdyn = 0;
        for (int b = 0; b < nb; ++b) {
            const BlockDesc& B = c.blocks[b];
            if (rho_class(B) == cls) {
                const size_t lds = rho_lds(B, pb);
                if (lds > (cls == 4 ? rho64_lds_max : (size_t)kMaxDynLds)) {
                    set_error("set_grid: a block with " + std::to_string(B.nao) + " AOs and " +
                              std::to_string(B.nshell) + " shells needs more LDS than a workgroup has "
                              "(raise ao_screen)");
                    return PDFXC_ERR_UNSUPPORTED;
                }
                rho_work.push_back(b);
                maxdyn = std::max(maxdyn, lds);
            }
        }
        r.count = (int)rho_work.size() - r.offset;
        r.lds = maxdyn;
        if (r.count) c.rho_launches.push_back(r);
    }
    // k_kmat work: lower tiles split into items of at most 96 tiles
    std::vector<KWork> kw[2];
    size_t klds[2] = {0, 0};
    for (int b = 0; b < nb; ++b) {
        const BlockDesc& B = c.blocks[b];
        const int T = B.ldn / 16;
        const int nt = T * (T + 1) / 2;
        const int cls = (nt <= 32) ? 0 : 1;
        // at most 8 x 12 tiles (the kernel's consumer limit); the tuning may cut items smaller
        const int per = (cls == 0) ? 32 : c.tune.k_tiles;
        for (int t0 = 0; t0 < nt; t0 += per) {
            KWork w;
            w.block = b;
            w.t0 = t0;
            w.t1 = std::min(nt, t0 + per);
            int I = 0;
            while ((I + 1) * (I + 2) / 2 <= w.t1 - 1) ++I;
            w.ncols = 16 * (I + 1);
            int nsh = 0;
            while (nsh < B.nshell && c.block_shell_col[B.shell_begin + nsh] < w.ncols) ++nsh;
            w.nsh = nsh;
            size_t lds = 0;
            const int LD = (w.ncols & 31) ? w.ncols : w.ncols + 16;
            for (int ps = 16; ps >= 4; ps /= 2) {
                w.ps = ps;
                lds = sizeof(double) * (4 * (size_t)ps * LD + (size_t)nsh * RS) +
                      sizeof(int) * (size_t)(w.ncols + 2 * nsh);
                if (lds <= (size_t)kMaxDynLds) break;
            }
            if (lds > (size_t)kMaxDynLds) {
                set_error("set_grid: block too wide for the K kernel's LDS");
                return PDFXC_ERR_UNSUPPORTED;
            }
            klds[cls] = std::max(klds[cls], lds);
            kw[cls].push_back(w);
        }
    }
    std::vector<KWork> kall;
    c.kmat_launches.clear();
    for (int cls = 0; cls < 2; ++cls) {
        if (kw[cls].empty()) continue;
        c.kmat_launches.push_back(LaunchRange{cls, (int)kall.size(), (int)kw[cls].size(), klds[cls]});
        kall.insert(kall.end(), kw[cls].begin(), kw[cls].end());
    }

    // split precision: chunked column layout and k_rho_split classes (ldnp <=
    // 256: 64 points per workgroup, <= 512: 32, <= 1024: 16); k_kmat_split items
    std::vector<int> srho_work;
    std::vector<KGWork> skw;
    size_t sk_lds = 0;
    c.sblocks.clear();
    c.chunks.clear();
    c.sshell_col.clear();
    c.srho_launches.clear();
    c.sk_launches.clear();
    c.k_fp64_launches.clear();
    c.n_k_fp64_blocks = 0;
    c.kgroups.clear();
    c.kg_blocks.clear();
    c.kg_shells.clear();
    c.kg_shell_col.clear();
    c.kg_sorder.clear();
    c.kg_cols.clear();
    c.kg_masks.clear();
    if (c.opts.precision != PDFXC_FP64) {
        const int S = prec_slices(c.opts.precision);
        c.sblocks.resize(nb);
        std::vector<int> ipos(c.n_aos);
        for (int q = 0; q < c.n_aos; ++q) ipos[c.split_perm[q]] = q;
        std::vector<int> ch;
        for (int b = 0; b < nb; ++b) {
            const BlockDesc& B = c.blocks[b];
            SplitBlock& SB = c.sblocks[b];
            SB.chunk_begin = (int)c.chunks.size();
            ch.clear();
            for (int k = 0; k < B.nshell; ++k) {
                const Shell& Sh = c.shells[c.block_shells[B.shell_begin + k]];
                const int q0 = ipos[Sh.first_ao];
                for (int g = q0 / 4; g <= (q0 + 2 * Sh.l) / 4; ++g) ch.push_back(g);
            }
            std::sort(ch.begin(), ch.end());
            ch.erase(std::unique(ch.begin(), ch.end()), ch.end());
            c.chunks.insert(c.chunks.end(), ch.begin(), ch.end());
            SB.nchunk = (int)ch.size();
            SB.ldnp = std::max(32, (4 * SB.nchunk + 31) / 32 * 32);
            SB.pad = 0;
            for (int k = 0; k < B.nshell; ++k) {
                const Shell& Sh = c.shells[c.block_shells[B.shell_begin + k]];
                const int q0 = ipos[Sh.first_ao];
                const int pos = (int)(std::lower_bound(ch.begin(), ch.end(), q0 / 4) - ch.begin());
                c.sshell_col.push_back(4 * pos + q0 % 4);
            }
            if (SB.ldnp > 1024 && prec_split_rho(c.opts.precision) && !prec_bfp(c.opts.precision)) {
                set_error("set_grid: a block keeps " + std::to_string(4 * SB.nchunk) +
                          " chunked AO columns; the split kernels take at most 1024 (raise ao_screen)");
                return PDFXC_ERR_UNSUPPORTED;
            }
        }
        auto srho_class = [&](int ldnp) {
            int base = 0;
            while ((256 << base) < ldnp) ++base;
            return std::min(2, base + c.tune.split_rho_shift);
        };
        // block floating point: classes 0, 1, 2 (64, 32, 16 points per
        // workgroup; blocks of at most 256, 384, 512 chunked columns), a block
        // in the first class that takes it and whose LDS fits, the tuning's
        // split_rho_shift moving blocks to narrower classes; blocks no class
        // takes run the FP64 k_rho, listed here as class kRhoFp64Class + their
        // k_rho class
        std::vector<int> bfp_cls;
        if (prec_bfp(c.opts.precision)) {
            size_t bfp_lds_max[3];
            for (int cls = 0; cls < 3; ++cls)
                if (int rc = bfp_max_dyn_lds(cls, &bfp_lds_max[cls])) return rc;
            auto bfp_lds = [&](int b, int cls) {
                const size_t ra = (size_t)c.sblocks[b].ldnp + 8;
                return std::max((size_t)kBfpPlanes * bfp_points_per_workgroup(cls) * ra * 2,
                                (size_t)4 * 512 * sizeof(double)) +
                       sizeof(double) * (size_t)c.blocks[b].nshell * RS +
                       sizeof(int) * (size_t)(c.sblocks[b].ldnp / 4 + 2 * c.blocks[b].nshell);
            };
            bfp_cls.assign(nb, -1);
            for (int b = 0; b < nb; ++b)
                for (int cls = std::min(2, c.tune.split_rho_shift); cls < 3; ++cls)
                    if (c.sblocks[b].ldnp <= bfp_max_columns(cls) && bfp_lds(b, cls) <= bfp_lds_max[cls]) {
                        bfp_cls[b] = cls;
                        break;
                    }
            for (int cls = 0; cls < 3; ++cls) {
                LaunchRange r{cls, (int)srho_work.size(), 0, 0};
                for (int b = 0; b < nb; ++b) {
                    if (bfp_cls[b] != cls) continue;
                    srho_work.push_back(b);
                    r.lds = std::max(r.lds, bfp_lds(b, cls));
                }
                r.count = (int)srho_work.size() - r.offset;
                if (r.count) c.srho_launches.push_back(r);
            }
        }
        for (int fcls = 0; fcls < 5 && prec_bfp(c.opts.precision); ++fcls) {
            LaunchRange r{kRhoFp64Class + fcls, (int)srho_work.size(), 0, 0};
            for (int b = 0; b < nb; ++b) {
                if (bfp_cls[b] >= 0 || rho_class(c.blocks[b]) != fcls) continue;
                srho_work.push_back(b);
                r.lds = std::max(r.lds, rho_lds(c.blocks[b], (size_t)rho_points_per_workgroup(fcls)));
            }
            r.count = (int)srho_work.size() - r.offset;
            if (r.count) c.srho_launches.push_back(r);
        }
        // (no split rho launches when the rho pass runs in FP64)
        for (int cls = 0; cls < 3 && prec_split_rho(c.opts.precision) && !prec_bfp(c.opts.precision); ++cls) {
            const size_t pb = (size_t)split_points_per_workgroup(cls);
            LaunchRange r{cls, (int)srho_work.size(), 0, 0};
            for (int b = 0; b < nb; ++b) {
                const SplitBlock& SB = c.sblocks[b];
                if (srho_class(SB.ldnp) != cls) continue;
                const size_t plane = std::max(((size_t)S * pb * (SB.ldnp + 8) * 2 + 15) / 16 * 16,
                                              (size_t)8 * 512 * sizeof(double));
                const size_t lds = plane + sizeof(double) * (size_t)c.blocks[b].nshell * RS +
                                   sizeof(int) * (size_t)(SB.ldnp / 4 + 2 * c.blocks[b].nshell);
                if (lds > (size_t)kMaxDynLds) {
                    set_error("set_grid: a block with " + std::to_string(c.blocks[b].nao) +
                              " AOs needs more LDS than a workgroup has in the split rho kernel (raise ao_screen)");
                    return PDFXC_ERR_UNSUPPORTED;
                }
                srho_work.push_back(b);
                r.lds = std::max(r.lds, lds);
            }
            r.count = (int)srho_work.size() - r.offset;
            if (r.count) c.srho_launches.push_back(r);
        }
        // k_kmat_split: groups of split_k_group neighbouring blocks (in the
        // bisection's leaf order, so siblings), summed on the union of their
        // shells; items of at most split_k_patches lower 2 x 2 tile patches in
        // column-major order; two slab buffers if they fit.  A group whose
        // widest item does not fit in LDS is halved; a single block that does
        // not fit takes the FP64 K pass.
        size_t k_lds_max = 0;
        if (int rc = split_kmat_max_dyn_lds(S, &k_lds_max)) return rc;
        std::vector<char> k_fp64_block(nb, 0);
        std::vector<int> block_of_leaf(nb);
        for (int li = 0; li < nb; ++li) block_of_leaf[lorder[li]] = li;
        const bool k_f16 = prec_k_f16(c.opts.precision);
        auto item_lds = [&](int wcols, int nsh_item, int nbuf) {
            return ((size_t)nbuf * 2 * S * wcols * 16 * 2 + 15) / 16 * 16 + sizeof(double) * (size_t)nsh_item * RS +
                   sizeof(int) * (size_t)(wcols + 3 * nsh_item + (k_f16 ? 2 * nbuf * wcols : 0));
        };
        struct GroupCost {
            double cost;
            int first, count;  // items in gitems
        };
        std::vector<KGWork> gitems;
        std::vector<GroupCost> gcost;
        std::vector<int> ushells;
        // union of the shells of blocks bl[0..n) (block shell lists ascend)
        auto make_union = [&](const int* bl, int n) {
            ushells.clear();
            for (int j = 0; j < n; ++j) {
                const BlockDesc& B = c.blocks[bl[j]];
                ushells.insert(ushells.end(), c.block_shells.begin() + B.shell_begin,
                               c.block_shells.begin() + B.shell_begin + B.nshell);
            }
            std::sort(ushells.begin(), ushells.end());
            ushells.erase(std::unique(ushells.begin(), ushells.end()), ushells.end());
        };
        // the first item (c0 = 0) is the widest: all union shells, all columns
        auto fits = [&](int nao, int nsh) {
            const int wcols = 32 * ((nao + 31) / 32);
            return nsh <= 32 * kMaxMaskWords && item_lds(wcols, nsh, 1) <= k_lds_max;
        };
        std::vector<int> bl;
        std::function<void(int, int)> form = [&](int leaf0, int n) {
            bl.clear();
            for (int j = 0; j < n; ++j) bl.push_back(block_of_leaf[leaf0 + j]);
            make_union(bl.data(), n);
            int nao = 0;
            for (int s : ushells) nao += 2 * c.shells[s].l + 1;
            if (!fits(nao, (int)ushells.size())) {
                if (n > 1) {
                    const int h = n / 2;
                    form(leaf0, h);
                    form(leaf0 + h, n - h);
                } else {
                    // the 16-point bf16 slabs of this block exceed a workgroup's
                    // LDS: its K pass runs in FP64 (k_kmat, whose slab depth
                    // adapts), which is at least as accurate
                    k_fp64_block[bl[0]] = 1;
                    ++c.n_k_fp64_blocks;
                }
                return;
            }
            KGroup G{};
            G.sub_begin = (int)c.kg_blocks.size();
            G.nsub = n;
            G.shell_begin = (int)c.kg_shells.size();
            G.nshell = (int)ushells.size();
            G.col_begin = (int)c.kg_cols.size();
            G.nao = nao;
            G.mask_begin = (int)c.kg_masks.size();
            G.nword = (G.nshell + 31) / 32;
            G.ldp = c.n_aos;
            G.pad = 0;
            G.poff = 0;
            int npts = 0;
            for (int j = 0; j < n; ++j) {
                c.kg_blocks.push_back(bl[j]);
                npts += c.blocks[bl[j]].npts;
            }
            int col = 0;
            std::vector<int> ucol(ushells.size());
            for (size_t k = 0; k < ushells.size(); ++k) {
                const Shell& Sh = c.shells[ushells[k]];
                c.kg_shells.push_back(ushells[k]);
                c.kg_shell_col.push_back(col);
                ucol[k] = col;
                for (int m = 0; m < 2 * Sh.l + 1; ++m) c.kg_cols.push_back(Sh.first_ao + m);
                col += 2 * Sh.l + 1;
            }
            for (int j = 0; j < n; ++j) {
                const BlockDesc& B = c.blocks[bl[j]];
                std::vector<unsigned> w(G.nword, 0u);
                for (int q = 0; q < B.nshell; ++q) {
                    const int k = (int)(std::lower_bound(ushells.begin(), ushells.end(),
                                                         c.block_shells[B.shell_begin + q]) - ushells.begin());
                    w[k >> 5] |= 1u << (k & 31);
                }
                c.kg_masks.insert(c.kg_masks.end(), w.begin(), w.end());
            }
            {
                // production order: equal (l, nprim) side by side, costliest last
                std::vector<int> so(ushells.size());
                std::iota(so.begin(), so.end(), 0);
                std::stable_sort(so.begin(), so.end(), [&](int x, int y) {
                    const Shell& A = c.shells[ushells[x]];
                    const Shell& Bs = c.shells[ushells[y]];
                    return A.l != Bs.l ? A.l < Bs.l : A.nprim < Bs.nprim;
                });
                c.kg_sorder.insert(c.kg_sorder.end(), so.begin(), so.end());
            }
            const int gi = (int)c.kgroups.size();
            c.kgroups.push_back(G);
            const int Tp = (nao + 31) / 32;
            const int np = Tp * (Tp + 1) / 2;
            const int per = c.tune.split_k_patches;  // at most 16, 24 or 32 (1, 2 or 4 per wave)
            GroupCost gc{0.0, (int)gitems.size(), 0};
            for (int q0 = 0; q0 < np; q0 += per) {
                KGWork w{};
                w.group = gi;
                w.q0 = q0;
                w.q1 = std::min(np, q0 + per);
                int Q = 0, q = q0, rem = Tp;
                while (q >= rem) {
                    q -= rem;
                    ++Q;
                    --rem;
                }
                w.c0 = 32 * Q;
                w.c1 = 32 * Tp;
                int k0 = 0;
                while (k0 < G.nshell && ucol[k0] + 2 * c.shells[ushells[k0]].l + 1 <= w.c0) ++k0;
                w.k0 = k0;
                w.nbuf = item_lds(w.c1 - w.c0, G.nshell - k0, 2) <= k_lds_max ? 2 : 1;
                const size_t lds = item_lds(w.c1 - w.c0, G.nshell - k0, w.nbuf);
                sk_lds = std::max(sk_lds, lds);
                w.pad = 0;
                gitems.push_back(w);
                // production (points x columns) plus MFMA (points x patches)
                gc.cost = std::max(gc.cost, (double)npts * ((w.c1 - w.c0) + 8.0 * (w.q1 - w.q0)));
            }
            gc.count = (int)gitems.size() - gc.first;
            gcost.push_back(gc);
        };
        const int G0 = c.tune.split_k_group;
        for (int leaf0 = 0; leaf0 < nb; leaf0 += G0) form(leaf0, std::min(G0, nb - leaf0));
        // launch order: groups by descending cost (their items side by side)
        std::vector<int> gorder(gcost.size());
        std::iota(gorder.begin(), gorder.end(), 0);
        std::stable_sort(gorder.begin(), gorder.end(),
                         [&](int x, int y) { return gcost[x].cost > gcost[y].cost; });
        for (int g : gorder)
            skw.insert(skw.end(), gitems.begin() + gcost[g].first, gitems.begin() + gcost[g].first + gcost[g].count);
        // launch class = workgroup shape for the item size (16 x 1, 12 x 2 or
        // 8 x 4 waves x patches), so a batch keeps every item's shape
        const int kshape = c.tune.split_k_patches <= 16 ? 0 : c.tune.split_k_patches <= 24 ? 1 : 2;
        if (!skw.empty()) c.sk_launches.push_back(LaunchRange{kshape, 0, (int)skw.size(), sk_lds});
        // FP64 K items of those blocks, appended to the FP64 work list
        for (int cls = 0; cls < 2; ++cls) {
            LaunchRange r{cls, (int)kall.size(), 0, klds[cls]};
            for (const KWork& w : kw[cls])
                if (k_fp64_block[w.block]) kall.push_back(w);
            r.count = (int)kall.size() - r.offset;
            if (r.count) c.k_fp64_launches.push_back(r);
        }
    }

    // device buffers
    free_grid_buffers(&c);
    int rc = dalloc(&c.d_pts, pts.size());
    if (rc == PDFXC_OK) rc = dalloc(&c.d_blocks, c.blocks.size());
    if (rc == PDFXC_OK) rc = dalloc(&c.d_block_shells, c.block_shells.size());
    if (rc == PDFXC_OK) rc = dalloc(&c.d_block_cols, c.block_cols.size());
    if (rc == PDFXC_OK) rc = dalloc(&c.d_block_shell_col, c.block_shell_col.size());
    if (rc == PDFXC_OK) rc = dalloc(&c.d_block_sorder, c.block_sorder.size());
    if (rc == PDFXC_OK) rc = dalloc(&c.d_rho_work, rho_work.size());
    if (rc == PDFXC_OK) rc = dalloc(reinterpret_cast<char**>(&c.d_k_work), kall.size() * sizeof(KWork));
    if (rc == PDFXC_OK) rc = dalloc(&c.d_coef, 4 * (size_t)c.n_slots);
    if (rc == PDFXC_OK) rc = dalloc(&c.d_rho, 4 * (size_t)c.n_slots);
    if (rc == PDFXC_OK) rc = dalloc(&c.d_eblock, (size_t)(c.n_slots + 255) / 256);
    if (rc == PDFXC_OK && c.opts.precision != PDFXC_FP64) {
        rc = dalloc(&c.d_sblocks, c.sblocks.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_chunks, c.chunks.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_sshell_col, c.sshell_col.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_srho_work, srho_work.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_kg_work, skw.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_kgroups, c.kgroups.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_kg_blocks, c.kg_blocks.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_kg_shells, c.kg_shells.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_kg_shell_col, c.kg_shell_col.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_kg_sorder, c.kg_sorder.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_kg_cols, c.kg_cols.size());
        if (rc == PDFXC_OK) rc = dalloc(&c.d_kg_masks, c.kg_masks.size());
    }
    if (rc != PDFXC_OK) return rc;
    auto up = [](void* d, const void* h, size_t n) { return hipMemcpy(d, h, n, hipMemcpyHostToDevice); };
    PDFXC_HIP(up(c.d_pts, pts.data(), pts.size() * sizeof(double)));
    PDFXC_HIP(up(c.d_blocks, c.blocks.data(), c.blocks.size() * sizeof(BlockDesc)));
    PDFXC_HIP(up(c.d_rho_work, rho_work.data(), rho_work.size() * sizeof(int)));
    PDFXC_HIP(up(c.d_k_work, kall.data(), kall.size() * sizeof(KWork)));
    if (!c.block_shells.empty()) {
        PDFXC_HIP(up(c.d_block_shells, c.block_shells.data(), c.block_shells.size() * sizeof(int)));
        PDFXC_HIP(up(c.d_block_shell_col, c.block_shell_col.data(), c.block_shell_col.size() * sizeof(int)));
        PDFXC_HIP(up(c.d_block_cols, c.block_cols.data(), c.block_cols.size() * sizeof(int)));
        PDFXC_HIP(up(c.d_block_sorder, c.block_sorder.data(), c.block_sorder.size() * sizeof(int)));
    }
    if (c.opts.precision != PDFXC_FP64) {
        PDFXC_HIP(up(c.d_sblocks, c.sblocks.data(), c.sblocks.size() * sizeof(SplitBlock)));
        PDFXC_HIP(up(c.d_chunks, c.chunks.data(), c.chunks.size() * sizeof(int)));
        PDFXC_HIP(up(c.d_sshell_col, c.sshell_col.data(), c.sshell_col.size() * sizeof(int)));
        PDFXC_HIP(up(c.d_srho_work, srho_work.data(), srho_work.size() * sizeof(int)));
        if (!skw.empty()) {
            PDFXC_HIP(up(c.d_kg_work, skw.data(), skw.size() * sizeof(KGWork)));
            PDFXC_HIP(up(c.d_kgroups, c.kgroups.data(), c.kgroups.size() * sizeof(KGroup)));
            PDFXC_HIP(up(c.d_kg_blocks, c.kg_blocks.data(), c.kg_blocks.size() * sizeof(int)));
            PDFXC_HIP(up(c.d_kg_shells, c.kg_shells.data(), c.kg_shells.size() * sizeof(int)));
            PDFXC_HIP(up(c.d_kg_shell_col, c.kg_shell_col.data(), c.kg_shell_col.size() * sizeof(int)));
            PDFXC_HIP(up(c.d_kg_sorder, c.kg_sorder.data(), c.kg_sorder.size() * sizeof(int)));
            PDFXC_HIP(up(c.d_kg_cols, c.kg_cols.data(), c.kg_cols.size() * sizeof(int)));
            PDFXC_HIP(up(c.d_kg_masks, c.kg_masks.data(), c.kg_masks.size() * sizeof(unsigned)));
        }
    }
    PDFXC_HIP(hipMemset(c.d_coef, 0, 4 * (size_t)c.n_slots * sizeof(double)));
    PDFXC_HIP(hipMemset(c.d_rho, 0, 4 * (size_t)c.n_slots * sizeof(double)));
    PDFXC_HIP(hipMemset(c.d_eblock, 0, (size_t)(c.n_slots + 255) / 256 * sizeof(double)));
    // host copies of the work lists (a batch concatenates them)
    c.h_rho_work = std::move(rho_work);
    c.h_k_work = std::move(kall);
    c.h_srho_work = std::move(srho_work);
    c.h_kg_work = std::move(skw);
    return PDFXC_OK;
}

// ---------------------------------------------------------------------------
// Grid generation pieces (DfGenerateGrid.cpp).  Lebedev rules are lists of
// octahedral orbits with the published Lebedev-Laikov parameters (the same
// rules as TlLebedevGrid.cpp:271-1577); each orbit is expanded by applying the
// 48 signed permutations to its seed point.
// ---------------------------------------------------------------------------
#include "lebedev_rules.inc"

struct Orbit {
    int kind;
    double a, b, w;
};

// supported orders, ascending (TlLebedevGrid.cpp:4-6)
static const int kLebOrders[] = {6,   14,  26,  38,  50,  74,  86,  110,  146,  170,  194,
                                 230, 266, 302, 350, 434, 590, 770, 974, 1202, 1454, 1730};

static bool lebedev_rule(int order, std::vector<Orbit>& rule) {
    const double(*t)[4] = nullptr;
    int no = 0;
#define LEB_CASE(N, T)                             \
    case N:                                        \
        t = T;                                     \
        no = (int)(sizeof(T) / sizeof(T[0]));      \
        break;
    switch (order) {
        LEBEDEV_RULE_LIST(LEB_CASE)
        default: return false;
    }
#undef LEB_CASE
    rule.clear();
    for (int k = 0; k < no; ++k) rule.push_back({(int)t[k][0], t[k][1], t[k][2], t[k][3]});
    return true;
}

// the largest supported order <= n, or 6 (getNumOfPrunedAnglarPoints, DfGenerateGrid.cpp:1677-1684)
static int lebedev_floor(int n) {
    int best = kLebOrders[0];
    for (int o : kLebOrders)
        if (o <= n) best = o;
    return best;
}

static bool lebedev_points(int order, std::vector<double>& xyz, std::vector<double>& w) {
    std::vector<Orbit> rule;
    if (!lebedev_rule(order, rule)) return false;
    xyz.clear();
    w.clear();
    for (const Orbit& o : rule) {
        double s[3];
        switch (o.kind) {
            case 1: s[0] = 1.0; s[1] = s[2] = 0.0; break;
            case 2: s[0] = s[1] = std::sqrt(0.5); s[2] = 0.0; break;
            case 3: s[0] = s[1] = s[2] = std::sqrt(1.0 / 3.0); break;
            case 4: s[0] = s[1] = o.a; s[2] = std::sqrt(1.0 - 2.0 * o.a * o.a); break;
            case 5: s[0] = o.a; s[1] = std::sqrt(1.0 - o.a * o.a); s[2] = 0.0; break;
            default: s[0] = o.a; s[1] = o.b; s[2] = std::sqrt(1.0 - o.a * o.a - o.b * o.b); break;
        }
        const size_t first = w.size();
        int perm[3] = {0, 1, 2};
        do {
            for (int sg = 0; sg < 8; ++sg) {
                double q[3];
                for (int d = 0; d < 3; ++d) {
                    const double v = s[perm[d]];
                    q[d] = ((sg >> d) & 1) ? -v : v;
                    if (q[d] == 0.0) q[d] = 0.0;
                }
                bool dup = false;
                for (size_t k = first; k < w.size() && !dup; ++k)
                    dup = xyz[3 * k] == q[0] && xyz[3 * k + 1] == q[1] && xyz[3 * k + 2] == q[2];
                if (!dup) {
                    xyz.insert(xyz.end(), q, q + 3);
                    w.push_back(o.w);
                }
            }
        } while (std::next_permutation(perm, perm + 3));
    }
    return (int)w.size() == order;
}

// Euler-Maclaurin radial abscissa and weight (DfGenerateGrid.cpp:1597-1610);
// i^5 is formed in 32-bit two's-complement as the reference does.
static void radial_em(double R, int Nr, int i, double* r, double* w) {
    const double inv = 1.0 / (double)(Nr + 1 - i);
    *r = R * (double)(i * i) * inv * inv;
    const uint32_t u = (uint32_t)i;
    const int32_t i5 = (int32_t)(u * u * u * u * u);
    *w = 2.0 * R * R * R * (double)(Nr + 1) * (double)i5 * inv * inv * inv * inv * inv * inv * inv *
         4.0 * kPi;
}

// Bragg-Slater radii as the grid uses them (TlPrdctbl.cpp:150-180)
static double bragg_slater(int Z) {
    static const double t[] = {0.00, 0.25, 0.50, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50,
                               0.50, 1.80, 1.50, 1.25, 1.10, 1.00, 1.00, 1.00, 1.00};
    return (Z >= 1 && Z < (int)(sizeof(t) / sizeof(t[0]))) ? t[Z] : 0.0;
}

// Gauss-Chebyshev (second kind) radial abscissa and weight with the Becke,
// TA or KK mapping (DfGenerateGrid.cpp:1612-1655); the weight carries 4 pi r^2.
static void radial_gc(double R, int Nr, int i, int mapping, double* r, double* w) {
    static const double inv_ln2 = 1.0 / std::log(2.0);
    const double Nr1 = Nr + 1;
    const double t = (double)i * kPi / Nr1;
    const double sn = std::sin(t), ss = sn * sn, cs = std::cos(t);
    const double x = (Nr1 - 2.0 * (double)i) / Nr1 + (2.0 / kPi) * (1.0 + (2.0 / 3.0) * ss) * cs * sn;
    const double inv_x1 = 1.0 / (x - 1.0);
    double ri, jac;
    if (mapping == PDFXC_GC_BECKE) {
        ri = R * (1.0 + x) / (1.0 - x);
        jac = R * ((x + 1.0) * inv_x1 * inv_x1 - inv_x1);
    } else if (mapping == PDFXC_GC_TA) {
        ri = inv_ln2 * std::log(2.0 * (-inv_x1)) * R;
        jac = R * inv_ln2 / (1.0 - x);
    } else {
        ri = inv_ln2 * std::log(2.0 * (-inv_x1));
        jac = inv_ln2 / (1.0 - x);
    }
    *r = ri;
    *w = (16.0 / (3.0 * Nr1)) * ss * ss * jac * ri * ri * 4.0 * kPi;
}

// SG-1 radial scales (setCellPara's SG-1 list, DfGenerateGrid.cpp:409-429)
static double sg1_radius(int Z) {
    static const double t[] = {0.0,    1.0000, 0.5882, 3.0769, 2.0513, 1.5385, 1.2308,
                               1.0256, 0.8791, 0.7692, 0.6838, 4.0909, 3.1579, 2.5714,
                               2.1687, 1.8750, 1.6514, 1.4754, 1.3333};
    return (Z >= 1 && Z < (int)(sizeof(t) / sizeof(t[0]))) ? t[Z] : 0.0;
}

// Lebedev order of the radial shell at r (DfGenerateGrid.cpp:1658-1713, the
// SG-1 regions :961-983): none -> n_ang floored to a supported order; MHL
// (K = 5, Bragg-Slater radius) -> min(int(5 r / R_B * n_ang), n_ang) floored;
// SG-1 -> 6, 38, 86, 194, 86 by r / R against the element row's alphas.
static int pruned_order(int pruning, int Z, double r, double R, int n_ang) {
    if (pruning == PDFXC_PRUNE_SG1) {
        double al[4] = {0.0, 0.0, 0.0, 10000.0};
        if (Z == 1 || Z == 2) {
            al[0] = 0.25, al[1] = 0.5, al[2] = 1.0, al[3] = 4.5;
        } else if (Z >= 3 && Z <= 10) {
            al[0] = 0.1667, al[1] = 0.5, al[2] = 0.9, al[3] = 3.5;
        } else if (Z >= 11 && Z <= 18) {
            al[0] = 0.1, al[1] = 0.4, al[2] = 0.8, al[3] = 2.5;
        }
        const double q = r * (1.0 / R);
        if (q < al[1]) return (q < al[0]) ? 6 : 38;
        if (q < al[2]) return 86;
        return (q < al[3]) ? 194 : 86;
    }
    int n = n_ang;
    if (pruning == PDFXC_PRUNE_MHL) {
        const double v = (5.0 * r / bragg_slater(Z)) * (double)n_ang;
        n = (v < (double)n_ang) ? (int)v : n_ang;
    }
    return lebedev_floor(n);
}

// Standard orientation (DfGenerateGrid.cpp:1739-1809): eigenvectors of the
// charge-weighted inertia-like tensor as columns, ascending eigenvalues; the
// largest-magnitude component of each column is made positive.
static void orientation(int na, const int* Z, const double* xyz, double O[9]) {
    double T[3] = {0, 0, 0}, sz = 0.0;
    for (int a = 0; a < na; ++a) {
        if (Z[a] <= 0) continue;
        for (int d = 0; d < 3; ++d) T[d] += Z[a] * xyz[3 * a + d];
        sz += Z[a];
    }
    for (int d = 0; d < 3; ++d) T[d] /= sz;
    double M[3][3] = {{0}};
    for (int a = 0; a < na; ++a) {
        if (Z[a] <= 0) continue;
        double r[3];
        for (int d = 0; d < 3; ++d) r[d] = xyz[3 * a + d] - T[d];
        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M[i][j] += Z[a] * ((i == j ? r2 : 0.0) - r[i] * r[j]);
    }
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = M[0][1] * M[0][1] + M[0][2] * M[0][2] + M[1][2] * M[1][2];
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (M[p][q] == 0.0) continue;
                const double th = (M[q][q] - M[p][p]) / (2.0 * M[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < 3; ++k) {
                    const double a = M[k][p], b = M[k][q];
                    M[k][p] = cs * a - sn * b;
                    M[k][q] = sn * a + cs * b;
                }
                for (int k = 0; k < 3; ++k) {
                    const double a = M[p][k], b = M[q][k];
                    M[p][k] = cs * a - sn * b;
                    M[q][k] = sn * a + cs * b;
                }
                for (int k = 0; k < 3; ++k) {
                    const double a = V[k][p], b = V[k][q];
                    V[k][p] = cs * a - sn * b;
                    V[k][q] = sn * a + cs * b;
                }
            }
    }
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int a, int b) { return M[a][a] < M[b][b]; });
    for (int col = 0; col < 3; ++col) {
        const int k = idx[col];
        int big = 0;
        for (int r = 1; r < 3; ++r)
            if (std::fabs(V[r][k]) > std::fabs(V[big][k])) big = r;
        const double sg = V[big][k] < 0 ? -1.0 : 1.0;
        for (int r = 0; r < 3; ++r) O[3 * r + col] = sg * V[r][k];
    }
}

int ensure_beta_buffers(Context* c) {
    const size_t n = 4 * (size_t)c->n_slots;
    if (!c->d_rho_b) {
        if (int rc = dalloc(&c->d_rho_b, n)) return rc;
        PDFXC_HIP(hipMemset(c->d_rho_b, 0, n * sizeof(double)));
    }
    if (!c->d_coef_b) {
        if (int rc = dalloc(&c->d_coef_b, n)) return rc;
        PDFXC_HIP(hipMemset(c->d_coef_b, 0, n * sizeof(double)));
    }
    return PDFXC_OK;
}

}  // namespace pdfxc

extern "C" {

int pdfxc_set_grid(pdfxc_ctx* ctx, int64_t n_pts, const double* x, const double* y, const double* z,
                   const double* w) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_set_grid: null context");
    if (n_pts <= 0 || !x || !y || !z || !w) return arg_error("pdfxc_set_grid: bad arguments");
    // a rejected grid leaves the context without one (a later build reports PDFXC_ERR_NO_GRID)
    if (c->stream) PDFXC_HIP(hipStreamSynchronize(c->stream));
    if (n_pts > ((int64_t)1 << 31) - kBlockPts) {
        reset_grid(*c);
        return arg_error("pdfxc_set_grid: more than 2^31 points");
    }
    for (int64_t i = 0; i < n_pts; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i]) || !std::isfinite(w[i])) {
            reset_grid(*c);
            return arg_error("pdfxc_set_grid: non-finite grid value at point " + std::to_string(i));
        }
    c->n_pts = n_pts;
    c->gen_atom.clear();
    c->grid_partition = PDFXC_PARTITION_BECKE;
    c->grid_size_adjust = 1;
    c->gx.assign(x, x + n_pts);
    c->gy.assign(y, y + n_pts);
    c->gz.assign(z, z + n_pts);
    c->gw.assign(w, w + n_pts);
    return build_blocks(*c);
}

int pdfxc_get_grid(pdfxc_ctx* ctx, double* x, double* y, double* z, double* w) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_get_grid: null context");
    if (c->n_pts == 0) {
        set_error("pdfxc_get_grid: no grid");
        return PDFXC_ERR_NO_GRID;
    }
    if (x) std::copy(c->gx.begin(), c->gx.end(), x);
    if (y) std::copy(c->gy.begin(), c->gy.end(), y);
    if (z) std::copy(c->gz.begin(), c->gz.end(), z);
    if (w) std::copy(c->gw.begin(), c->gw.end(), w);
    return PDFXC_OK;
}

int pdfxc_grid_opts_preset(const char* grid_type, pdfxc_grid_opts* out) {
    if (!grid_type || !out) return arg_error("pdfxc_grid_opts_preset: bad arguments");
    std::string t(grid_type);
    for (char& ch : t) ch = (ch == '_') ? '-' : (char)std::toupper((unsigned char)ch);
    pdfxc_grid_opts o;
    o.n_rad = 75;
    o.n_ang = 302;
    o.radial = PDFXC_RADIAL_EM;
    o.gc_mapping = PDFXC_GC_TA;
    o.pruning = PDFXC_PRUNE_MHL;
    o.partitioning = PDFXC_PARTITION_SS;
    o.size_adjust = 1;
    o.radii = PDFXC_RADII_BRAGG;
    o.weight_cutoff = 1.0e-16;
    if (t == "COARSE") {
        o.n_rad = 35, o.n_ang = 110;
    } else if (t == "MEDIUM") {
        o.n_rad = 32, o.n_ang = 146;
    } else if (t == "MEDIUM-FINE") {
        o.n_rad = 64, o.n_ang = 146;
    } else if (t == "FINE" || t == "USER") {
        o.n_rad = 75, o.n_ang = 302;
    } else if (t == "ULTRA-FINE") {
        o.n_rad = 99, o.n_ang = 590;
    } else if (t == "SG-1" || t == "SG1") {
        o.n_rad = 50, o.n_ang = 194;
        o.pruning = PDFXC_PRUNE_SG1;
        o.radii = PDFXC_RADII_SG1;
        o.size_adjust = 0;
    } else {
        set_error(std::string("pdfxc_grid_opts_preset: unsupported grid type: ") + grid_type);
        return PDFXC_ERR_UNSUPPORTED;
    }
    *out = o;
    return PDFXC_OK;
}

static int generate_grid_impl(pdfxc_ctx* ctx, int n_atoms, const int* Z, const double* xyz,
                              const pdfxc_grid_opts* opts, const double* O_given, int64_t* n_out);

int pdfxc_generate_grid_ex(pdfxc_ctx* ctx, int n_atoms, const int* Z, const double* xyz,
                           const pdfxc_grid_opts* opts, int64_t* n_out) {
    return generate_grid_impl(ctx, n_atoms, Z, xyz, opts, nullptr, n_out);
}

int pdfxc_generate_grid_oriented(pdfxc_ctx* ctx, int n_atoms, const int* Z, const double* xyz,
                                 const pdfxc_grid_opts* opts, const double* O, int64_t* n_out) {
    if (!O) return arg_error("pdfxc_generate_grid_oriented: null orientation");
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double d = 0.0;  // O^T O = 1 (columns orthonormal)
            for (int k = 0; k < 3; ++k) d += O[3 * k + i] * O[3 * k + j];
            if (!(std::fabs(d - (i == j ? 1.0 : 0.0)) <= 1e-10))
                return arg_error("pdfxc_generate_grid_oriented: the orientation is not orthogonal");
        }
    return generate_grid_impl(ctx, n_atoms, Z, xyz, opts, O, n_out);
}

static int generate_grid_impl(pdfxc_ctx* ctx, int n_atoms, const int* Z, const double* xyz,
                              const pdfxc_grid_opts* opts, const double* O_given, int64_t* n_out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_generate_grid_ex: null context");
    if (n_atoms <= 0 || !Z || !xyz || !opts) return arg_error("pdfxc_generate_grid_ex: bad arguments");
    const pdfxc_grid_opts o = *opts;
    if (o.n_rad <= 0 || o.n_ang <= 0 || (o.radial != PDFXC_RADIAL_EM && o.radial != PDFXC_RADIAL_GC) ||
        o.gc_mapping < PDFXC_GC_BECKE || o.gc_mapping > PDFXC_GC_KK || o.pruning < PDFXC_PRUNE_NONE ||
        o.pruning > PDFXC_PRUNE_SG1 || (o.partitioning != PDFXC_PARTITION_BECKE && o.partitioning != PDFXC_PARTITION_SS) ||
        (o.radii != PDFXC_RADII_BRAGG && o.radii != PDFXC_RADII_SG1))
        return arg_error("pdfxc_generate_grid_ex: bad options");
    for (int a = 0; a < n_atoms; ++a)
        if (Z[a] < 0 || (Z[a] > 0 && bragg_slater(Z[a]) == 0.0)) {
            set_error("pdfxc_generate_grid_ex: no Bragg-Slater radius for Z = " + std::to_string(Z[a]));
            return PDFXC_ERR_UNSUPPORTED;
        }
    // Lebedev rules a shell can take: every supported order up to the largest
    const int top = (o.pruning == PDFXC_PRUNE_SG1) ? 194 : lebedev_floor(o.n_ang);
    std::vector<std::vector<double>> leb(sizeof(kLebOrders) / sizeof(kLebOrders[0])),
        lebw(sizeof(kLebOrders) / sizeof(kLebOrders[0]));
    auto rule_index = [](int order) {
        int q = 0;
        while (kLebOrders[q] != order) ++q;
        return q;
    };
    for (int q = 0; q < (int)leb.size() && kLebOrders[q] <= top; ++q)
        if (!lebedev_points(kLebOrders[q], leb[q], lebw[q])) {
            set_error("pdfxc_generate_grid_ex: Lebedev rule " + std::to_string(kLebOrders[q]) + " is inconsistent");
            return PDFXC_ERR_UNSUPPORTED;
        }
    double O[9];
    if (O_given)
        std::copy(O_given, O_given + 9, O);
    else
        orientation(n_atoms, Z, xyz, O);
    // single-centre points and weights, atom by atom, radial shell by shell
    std::vector<double> pts, w;
    std::vector<int> atom;
    for (int a = 0; a < n_atoms; ++a) {
        if (Z[a] <= 0) continue;
        const double R = (o.radii == PDFXC_RADII_SG1) ? sg1_radius(Z[a]) : bragg_slater(Z[a]);
        for (int i = 1; i <= o.n_rad; ++i) {
            double ri, wr;
            if (o.radial == PDFXC_RADIAL_GC)
                radial_gc(R, o.n_rad, i, o.gc_mapping, &ri, &wr);
            else
                radial_em(R, o.n_rad, i, &ri, &wr);
            if (ri > 30.0) continue;
            const int q = rule_index(pruned_order(o.pruning, Z[a], ri, R, o.n_ang));
            const std::vector<double>& L = leb[q];
            const int nq = (int)lebw[q].size();
            for (int k = 0; k < nq; ++k) {
                for (int d = 0; d < 3; ++d) {
                    const double v = O[3 * d] * L[3 * k] + O[3 * d + 1] * L[3 * k + 1] + O[3 * d + 2] * L[3 * k + 2];
                    pts.push_back(v * ri + xyz[3 * a + d]);
                }
                w.push_back(lebw[q][k] * wr);
                atom.push_back(a);
            }
        }
    }
    const int64_t n = (int64_t)w.size();
    // atom table for the partition kernel: (x, y, z, Z), 1/R_mn, Becke size
    // adjustments a_mn (getCovalentRadiiForBecke: H 0.35, else Bragg-Slater,
    // a dummy 0; |a| <= 1/2; zero when the adjustments are off)
    std::vector<double> at(4 * (size_t)n_atoms + 2 * (size_t)n_atoms * n_atoms, 0.0);
    for (int m = 0; m < n_atoms; ++m) {
        for (int d = 0; d < 3; ++d) at[4 * m + d] = xyz[3 * m + d];
        at[4 * m + 3] = Z[m];
    }
    double* invd = at.data() + 4 * n_atoms;
    double* amat = invd + (size_t)n_atoms * n_atoms;
    auto brad = [](int z) { return z == 1 ? 0.35 : bragg_slater(z); };
    for (int m = 0; m < n_atoms; ++m)
        for (int k = 0; k < n_atoms; ++k) {
            if (m == k) continue;
            const double dx = xyz[3 * m] - xyz[3 * k], dy = xyz[3 * m + 1] - xyz[3 * k + 1],
                         dz = xyz[3 * m + 2] - xyz[3 * k + 2];
            const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (dist < 1e-10)
                return arg_error("pdfxc_generate_grid_ex: atoms " + std::to_string(m) + " and " +
                                 std::to_string(k) + " coincide");
            invd[(size_t)m * n_atoms + k] = 1.0 / dist;
            if (o.size_adjust && Z[k] > 0) {
                // a dummy m has R = 0: au = -1, a = -1/0 clamps to -1/2 as in the reference
                const double Rm = Z[m] > 0 ? brad(Z[m]) : 0.0, Rk = brad(Z[k]);
                const double au = (Rm - Rk) / (Rm + Rk);
                const double den = au * au - 1.0;
                double aa = (den == 0.0) ? (au > 0.0 ? 0.5 : -0.5) : au / den;
                aa = std::min(0.5, std::max(-0.5, aa));
                amat[(size_t)k * n_atoms + m] = aa;  // a_mk at [k][m]: the kernel's lanes run over m
            }
        }
    double *d_at = nullptr, *d_p = nullptr, *d_w = nullptr;
    int* d_a = nullptr;
    int rc = dalloc(&d_at, at.size());
    if (rc == PDFXC_OK && n > 0) rc = dalloc(&d_p, pts.size());
    if (rc == PDFXC_OK && n > 0) rc = dalloc(&d_w, w.size());
    if (rc == PDFXC_OK && n > 0) rc = dalloc(&d_a, atom.size());
    hipError_t e = hipSuccess;
    if (rc == PDFXC_OK && n > 0) {
        e = hipMemcpyAsync(d_at, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_p, pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_a, atom.data(), atom.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_generate_grid_ex: upload");
        if (rc == PDFXC_OK) rc = launch_becke(d_at, n_atoms, n, d_p, d_a, d_w, o.partitioning, c->stream);
        if (rc == PDFXC_OK) {
            e = hipMemcpyAsync(w.data(), d_w, w.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = hip_fail(e, "pdfxc_generate_grid_ex: partition weights");
        }
    }
    dfree(d_at);
    dfree(d_p);
    dfree(d_w);
    dfree(d_a);
    if (rc != PDFXC_OK) return rc;
    // weight screening (screeningGridsByWeight0, DfGenerateGrid.cpp:1979-2004)
    std::vector<double> X, Y, Zc, W;
    std::vector<int> kept_atom;
    for (int64_t i = 0; i < n; ++i)
        if (std::fabs(w[i]) > o.weight_cutoff) {
            X.push_back(pts[3 * i]);
            Y.push_back(pts[3 * i + 1]);
            Zc.push_back(pts[3 * i + 2]);
            W.push_back(w[i]);
            kept_atom.push_back(atom[i]);
        }
    if (W.empty()) return arg_error("pdfxc_generate_grid_ex: no grid point survives the weight screen");
    if (n_out) *n_out = (int64_t)W.size();
    rc = pdfxc_set_grid(ctx, (int64_t)W.size(), X.data(), Y.data(), Zc.data(), W.data());
    if (rc == PDFXC_OK) {
        c->gen_atom = std::move(kept_atom);
        c->grid_partition = o.partitioning;
        c->grid_size_adjust = o.size_adjust ? 1 : 0;
    }
    return rc;
}

int pdfxc_generate_grid(pdfxc_ctx* ctx, int n_atoms, const int* Z, const double* xyz, int n_rad,
                        int n_ang, double weight_cutoff, int64_t* n_out) {
    if (!ctx) return arg_error("pdfxc_generate_grid: null context");
    if (lebedev_floor(n_ang) != n_ang) {
        set_error("pdfxc_generate_grid: unsupported Lebedev order " + std::to_string(n_ang) +
                  " (6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590, 770, 974, "
                  "1202, 1454, 1730)");
        return PDFXC_ERR_UNSUPPORTED;
    }
    pdfxc_grid_opts o;
    o.n_rad = n_rad;
    o.n_ang = n_ang;
    o.radial = PDFXC_RADIAL_EM;
    o.gc_mapping = PDFXC_GC_TA;
    o.pruning = PDFXC_PRUNE_NONE;
    o.partitioning = PDFXC_PARTITION_BECKE;
    o.size_adjust = 1;
    o.radii = PDFXC_RADII_BRAGG;
    o.weight_cutoff = weight_cutoff;
    return pdfxc_generate_grid_ex(ctx, n_atoms, Z, xyz, &o, n_out);
}

// Functional ids are the XC_TYPE values of DfXCFunctional.h:44; an unknown
// one is refused as the reference's getFxc aborts on it (DfXCFunctional.h:182-186).
static int check_functional(int functional, const char* who) {
    if (xc_type_valid(functional)) return PDFXC_OK;
    set_error(std::string(who) + ": unknown functional " + std::to_string(functional) +
              " (1 HF, 2 HFS, 3 SVWN, 4 HFB, 5 BLYP, 6 B3LYP)");
    return PDFXC_ERR_UNSUPPORTED;
}

// rho pass of one density into rho (the context's buffer, or the beta one)
static int run_rho(Context* c, const double* dP, double* rho) {
    double* saveP = c->d_P;
    double* saveR = c->d_rho;
    c->d_P = const_cast<double*>(dP);
    c->d_rho = rho;
    int rc = PDFXC_OK;
    if (prec_split_rho(c->opts.precision)) {
        rc = launch_split_rho(*c);
    } else {
        for (const LaunchRange& r : c->rho_launches)
            if (rc == PDFXC_OK) rc = launch_rho(*c, r);
    }
    c->d_P = saveP;
    c->d_rho = saveR;
    return rc;
}

// K pass of one set of point terms into dK (lower triangle accumulated)
static int run_kmat(Context* c, double* coef, double* dK) {
    double* saveC = c->d_coef;
    double* saveK = c->d_K;
    c->d_coef = coef;
    c->d_K = dK;
    int rc = PDFXC_OK;
    if (c->opts.precision != PDFXC_FP64) {
        rc = launch_split_kmat(*c);
        for (const LaunchRange& r : c->k_fp64_launches)
            if (rc == PDFXC_OK) rc = launch_kmat(*c, r);
    } else {
        for (const LaunchRange& r : c->kmat_launches)
            if (rc == PDFXC_OK) rc = launch_kmat(*c, r);
    }
    c->d_coef = saveC;
    c->d_K = saveK;
    return rc;
}

// Closes a build sequence's work-queue counters however the sequence ends, so
// that launches outside a sequence never take a counter begin_queues did not zero.
struct QueueScope {
    Context* c;
    explicit QueueScope(Context* cc) : c(cc) {}
    ~QueueScope() { c->queue_open = false; }
};

// One RKS (nspin 1) or UKS (nspin 2) build on device buffers: the rho pass of
// each spin, the point terms, the K pass of each spin, the energy sum.
static int build_device(Context* c, int xc, int nspin, const double* const* dP, double* const* dK, double* dE,
                        int accumulate_rho) {
    if (c->n_pts == 0 || c->blocks.empty()) {
        set_error("build: no grid set");
        return PDFXC_ERR_NO_GRID;
    }
    if (xc == kHF) {
        // HF has no grid term: E_xc = 0 and K untouched (DfXCFunctional.h:141-143)
        const int rc = launch_zero(dE, 1, c->stream);
        if (rc) return rc;
        PDFXC_HIP(hipStreamSynchronize(c->stream));
        for (double& t : c->last_kernel_ms) t = 0.0;
        c->last_build_ms = 0.0;
        return PDFXC_OK;
    }
    if (nspin == 2)
        if (int rc = ensure_beta_buffers(c)) return rc;
    QueueScope queues(c);
    c->rho_accumulate = accumulate_rho;
    PDFXC_HIP(hipEventRecord(c->ev[0], c->stream));
    int rc = begin_queues(*c);
    if (rc == PDFXC_OK) rc = run_rho(c, dP[0], c->d_rho);
    if (nspin == 2 && rc == PDFXC_OK) rc = run_rho(c, dP[1], c->d_rho_b);
    if (rc == PDFXC_OK) rc = nspin == 2 ? launch_xc_points_uks(*c, xc) : launch_xc_points(*c, xc);
    PDFXC_HIP(hipEventRecord(c->ev[1], c->stream));
    if (rc == PDFXC_OK) rc = run_kmat(c, c->d_coef, dK[0]);
    if (nspin == 2 && rc == PDFXC_OK) rc = run_kmat(c, c->d_coef_b, dK[1]);
    PDFXC_HIP(hipEventRecord(c->ev[2], c->stream));
    if (rc == PDFXC_OK) rc = launch_energy_sum(*c, dE);
    PDFXC_HIP(hipEventRecord(c->ev[3], c->stream));
    c->rho_accumulate = 0;
    if (rc) return rc;
    PDFXC_HIP(hipEventSynchronize(c->ev[3]));
    float t1, t2, t3;
    PDFXC_HIP(hipEventElapsedTime(&t1, c->ev[0], c->ev[1]));
    PDFXC_HIP(hipEventElapsedTime(&t2, c->ev[1], c->ev[2]));
    PDFXC_HIP(hipEventElapsedTime(&t3, c->ev[2], c->ev[3]));
    c->last_kernel_ms[0] = t1;
    c->last_kernel_ms[1] = t2;
    c->last_kernel_ms[2] = t3;
    c->last_kernel_ms[3] = 0.0;
    c->last_build_ms = t1 + t2 + t3;
    return PDFXC_OK;
}

static int build_device(Context* c, int xc, const double* dP, double* dK, double* dE, int accumulate_rho = 0) {
    const double* P[1] = {dP};
    double* K[1] = {dK};
    return build_device(c, xc, 1, P, K, dE, accumulate_rho);
}

int pdfxc_build_rks_device(pdfxc_ctx* ctx, int functional, const double* dP, double* dK, double* dE) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_build_rks_device: null context");
    if (int rc = check_functional(functional, "pdfxc_build_rks_device")) return rc;
    if (!dP || !dK || !dE) return arg_error("pdfxc_build_rks_device: null buffer");
    return build_device(c, functional, dP, dK, dE);
}

int pdfxc_build_rks_update_device(pdfxc_ctx* ctx, int functional, const double* d_dP, double* dK, double* dE) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_build_rks_update_device: null context");
    if (int rc = check_functional(functional, "pdfxc_build_rks_update_device")) return rc;
    if (!d_dP || !dK || !dE) return arg_error("pdfxc_build_rks_update_device: null buffer");
    return build_device(c, functional, d_dP, dK, dE, 1);
}

static int build_host(Context* c, int xc, const double* P_packed, double* K_packed, double* E_xc,
                      double* rho_out, int accumulate_rho);
static int build_device(Context* c, int xc, int nspin, const double* const* dP, double* const* dK, double* dE,
                        int accumulate_rho);

int pdfxc_build_rks_update(pdfxc_ctx* ctx, int functional, const double* dP_packed, double* K_packed,
                           double* E_xc, double* rho_out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_build_rks_update: null context");
    if (int rc = check_functional(functional, "pdfxc_build_rks_update")) return rc;
    if (!dP_packed || !K_packed || !E_xc) return arg_error("pdfxc_build_rks_update: null buffer");
    return build_host(c, functional, dP_packed, K_packed, E_xc, rho_out, 1);
}

int pdfxc_build_rks(pdfxc_ctx* ctx, int functional, const double* P_packed, double* K_packed,
                    double* E_xc, double* rho_out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_build_rks: null context");
    if (int rc = check_functional(functional, "pdfxc_build_rks")) return rc;
    if (!P_packed || !K_packed || !E_xc) return arg_error("pdfxc_build_rks: null buffer");
    return build_host(c, functional, P_packed, K_packed, E_xc, rho_out, 0);
}

static int build_host(Context* c, int xc, int nspin, const double* const* P_packed, double* const* K_packed,
                      double* E_xc, double* rho_out, int accumulate_rho) {
    if (c->n_pts == 0 || c->blocks.empty()) {
        set_error("build: no grid set");
        return PDFXC_ERR_NO_GRID;
    }
    if (xc == kHF) {
        // no grid term: E_xc = 0, K and rho_out untouched (DfXCFunctional.h:141-143)
        *E_xc = 0.0;
        for (double& t : c->last_kernel_ms) t = 0.0;
        c->last_build_ms = 0.0;
        return PDFXC_OK;
    }
    const int N = c->n_aos;
    const size_t NN = (size_t)N * N;
    if (nspin == 2 && !c->d_P2) {
        if (int rc = dalloc(&c->d_P2, NN)) return rc;
        if (int rc = dalloc(&c->d_K2, NN)) return rc;
    }
    double* dP[2] = {c->d_P, c->d_P2};
    double* dK[2] = {c->d_K, c->d_K2};
    double* H = c->h_pinned;
    for (int sp = 0; sp < nspin; ++sp) {
        for (int r = 0, k = 0; r < N; ++r)
            for (int q = 0; q <= r; ++q, ++k) {
                H[(size_t)r * N + q] = P_packed[sp][k];
                H[(size_t)q * N + r] = P_packed[sp][k];
            }
        PDFXC_HIP(hipMemcpyAsync(dP[sp], H, sizeof(double) * NN, hipMemcpyHostToDevice, c->stream));
        PDFXC_HIP(hipMemsetAsync(dK[sp], 0, sizeof(double) * NN, c->stream));
        PDFXC_HIP(hipStreamSynchronize(c->stream));  // H is reused
    }
    int rc = build_device(c, xc, nspin, dP, dK, c->d_E, accumulate_rho);
    if (rc) return rc;
    double E = 0.0;
    PDFXC_HIP(hipMemcpyAsync(&E, c->d_E, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    for (int sp = 0; sp < nspin; ++sp) {
        PDFXC_HIP(hipMemcpyAsync(H, dK[sp], sizeof(double) * NN, hipMemcpyDeviceToHost, c->stream));
        PDFXC_HIP(hipStreamSynchronize(c->stream));
        for (int r = 0, k = 0; r < N; ++r)
            for (int q = 0; q <= r; ++q, ++k) K_packed[sp][k] += H[(size_t)r * N + q];
    }
    *E_xc = E;
    if (rho_out) {
        const int stride = 4 * nspin;
        std::vector<double> rs(4 * (size_t)c->n_slots);
        for (int sp = 0; sp < nspin; ++sp) {
            PDFXC_HIP(hipMemcpy(rs.data(), sp ? c->d_rho_b : c->d_rho, rs.size() * sizeof(double),
                                hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < c->n_pts; ++p) {
                const int64_t sl = c->slot_of_point[p];
                for (int q = 0; q < 4; ++q) rho_out[stride * p + 4 * sp + q] = rs[4 * sl + q];
            }
        }
    }
    return PDFXC_OK;
}

static int build_host(Context* c, int xc, const double* P_packed, double* K_packed, double* E_xc,
                      double* rho_out, int accumulate_rho) {
    const double* P[1] = {P_packed};
    double* K[1] = {K_packed};
    return build_host(c, xc, 1, P, K, E_xc, rho_out, accumulate_rho);
}

// UKS (calcXCIntegForFockAndEnergy(P_A, P_B, ...), DfCalcGridX.cpp:2291-2330)
static int uks_args(Context* c, int functional, const void* a, const void* b, const void* ka, const void* kb,
                    const void* e, const char* who) {
    if (!c) return arg_error(std::string(who) + ": null context");
    if (int rc = check_functional(functional, who)) return rc;
    if (!a || !b || !ka || !kb || !e) return arg_error(std::string(who) + ": null buffer");
    return PDFXC_OK;
}

int pdfxc_build_uks(pdfxc_ctx* ctx, int functional, const double* PA_packed, const double* PB_packed,
                    double* KA_packed, double* KB_packed, double* E_xc, double* rho_out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (int rc = uks_args(c, functional, PA_packed, PB_packed, KA_packed, KB_packed, E_xc, "pdfxc_build_uks"))
        return rc;
    const double* P[2] = {PA_packed, PB_packed};
    double* K[2] = {KA_packed, KB_packed};
    return build_host(c, functional, 2, P, K, E_xc, rho_out, 0);
}

int pdfxc_build_uks_update(pdfxc_ctx* ctx, int functional, const double* dPA_packed, const double* dPB_packed,
                           double* KA_packed, double* KB_packed, double* E_xc, double* rho_out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (int rc = uks_args(c, functional, dPA_packed, dPB_packed, KA_packed, KB_packed, E_xc,
                          "pdfxc_build_uks_update"))
        return rc;
    const double* P[2] = {dPA_packed, dPB_packed};
    double* K[2] = {KA_packed, KB_packed};
    return build_host(c, functional, 2, P, K, E_xc, rho_out, 1);
}

int pdfxc_build_uks_device(pdfxc_ctx* ctx, int functional, const double* dPA, const double* dPB, double* dKA,
                           double* dKB, double* dE) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (int rc = uks_args(c, functional, dPA, dPB, dKA, dKB, dE, "pdfxc_build_uks_device")) return rc;
    const double* P[2] = {dPA, dPB};
    double* K[2] = {dKA, dKB};
    return build_device(c, functional, 2, P, K, dE, 0);
}

int pdfxc_build_uks_update_device(pdfxc_ctx* ctx, int functional, const double* d_dPA, const double* d_dPB,
                                  double* dKA, double* dKB, double* dE) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (int rc = uks_args(c, functional, d_dPA, d_dPB, dKA, dKB, dE, "pdfxc_build_uks_update_device")) return rc;
    const double* P[2] = {d_dPA, d_dPB};
    double* K[2] = {dKA, dKB};
    return build_device(c, functional, 2, P, K, dE, 1);
}

int pdfxc_eval_functional_uks(int functional, int64_t n, const double* in, double* out) {
    if (int rc = check_functional(functional, "pdfxc_eval_functional_uks")) return rc;
    if (n < 0 || (n > 0 && (!in || !out))) return arg_error("pdfxc_eval_functional_uks: bad arguments");
    if (n == 0) return PDFXC_OK;
    if (functional == kHF) {
        std::fill(out, out + 6 * n, 0.0);
        return PDFXC_OK;
    }
    double *di = nullptr, *dout = nullptr;
    int rc = dalloc(&di, 5 * (size_t)n);
    if (rc == PDFXC_OK) rc = dalloc(&dout, 6 * (size_t)n);
    if (rc == PDFXC_OK) {
        hipError_t e = hipMemcpy(di, in, 5 * n * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_eval_functional_uks: upload");
    }
    if (rc == PDFXC_OK) rc = launch_eval_functional_uks(functional, n, di, dout, nullptr);
    if (rc == PDFXC_OK) {
        hipError_t e = hipMemcpy(out, dout, 6 * n * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_eval_functional_uks: download");
    }
    dfree(di);
    dfree(dout);
    return rc;
}

int pdfxc_gradient_rks(pdfxc_ctx* ctx, int functional, int n_atoms, const int* Z, const double* xyz,
                       const int* point_atom, const double* P_packed, double* grad_functional,
                       double* grad_weight, double* E_xc) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_gradient_rks: null context");
    if (int rc = check_functional(functional, "pdfxc_gradient_rks")) return rc;
    if (n_atoms <= 0 || !Z || !xyz || !P_packed || !grad_functional || !grad_weight)
        return arg_error("pdfxc_gradient_rks: bad arguments");
    if (c->n_pts == 0 || c->blocks.empty()) {
        set_error("pdfxc_gradient_rks: no grid set");
        return PDFXC_ERR_NO_GRID;
    }
    if (!point_atom && c->gen_atom.size() != (size_t)c->n_pts)
        return arg_error("pdfxc_gradient_rks: the grid was not made by pdfxc_generate_grid; pass point_atom");
    for (int a = 0; a < n_atoms; ++a)
        if (Z[a] <= 0) {
            set_error("pdfxc_gradient_rks: dummy atoms (Z = 0) are not supported");
            return PDFXC_ERR_UNSUPPORTED;
        }
    const int N = c->n_aos;
    const int na = n_atoms;
    if (functional == kHF) {
        // no grid term, so no XC gradient
        std::fill(grad_functional, grad_functional + 3 * na, 0.0);
        std::fill(grad_weight, grad_weight + 3 * na, 0.0);
        if (E_xc) *E_xc = 0.0;
        return PDFXC_OK;
    }
    // atom of every shell: the atom at its centre
    std::vector<int> shell_atom(c->shells.size(), -1);
    for (size_t k = 0; k < c->shells.size(); ++k) {
        const double* ctr = c->shells[k].center;
        for (int a = 0; a < na; ++a) {
            const double dx = ctr[0] - xyz[3 * a], dy = ctr[1] - xyz[3 * a + 1], dz = ctr[2] - xyz[3 * a + 2];
            if (dx * dx + dy * dy + dz * dz < 1e-16) {
                shell_atom[k] = a;
                break;
            }
        }
        if (shell_atom[k] < 0)
            return arg_error("pdfxc_gradient_rks: the centre of AO " + std::to_string(c->shells[k].first_ao) +
                             " is not at any atom");
    }
    std::vector<int> slot_atom((size_t)c->n_slots, -1);
    for (int64_t p = 0; p < c->n_pts; ++p) {
        const int a = point_atom ? point_atom[p] : c->gen_atom[p];
        if (a < 0 || a >= na || Z[a] <= 0)
            return arg_error("pdfxc_gradient_rks: point " + std::to_string(p) + " has no valid atom");
        slot_atom[c->slot_of_point[p]] = a;
    }
    // Becke table (x, y, z, Z | 1/R_mn | a_mn) as the grid generator builds it
    std::vector<double> at(4 * (size_t)na + 2 * (size_t)na * na, 0.0);
    for (int m = 0; m < na; ++m) {
        for (int d = 0; d < 3; ++d) at[4 * m + d] = xyz[3 * m + d];
        at[4 * m + 3] = Z[m];
    }
    {
        double* invd = at.data() + 4 * na;
        double* amat = invd + (size_t)na * na;
        auto brad = [](int z) { return z == 1 ? 0.35 : bragg_slater(z); };
        for (int m = 0; m < na; ++m)
            for (int k = 0; k < na; ++k) {
                if (m == k) continue;
                const double dx = xyz[3 * m] - xyz[3 * k], dy = xyz[3 * m + 1] - xyz[3 * k + 1],
                             dz = xyz[3 * m + 2] - xyz[3 * k + 2];
                const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (dist < 1e-10 && Z[m] > 0 && Z[k] > 0)
                    return arg_error("pdfxc_gradient_rks: atoms " + std::to_string(m) + " and " +
                                     std::to_string(k) + " coincide");
                invd[(size_t)m * na + k] = 1.0 / dist;
                if (Z[m] > 0 && Z[k] > 0) {
                    if (bragg_slater(Z[m]) == 0.0 || bragg_slater(Z[k]) == 0.0) {
                        set_error("pdfxc_gradient_rks: no Bragg-Slater radius for an atom");
                        return PDFXC_ERR_UNSUPPORTED;
                    }
                    const double Rm = brad(Z[m]), Rk = brad(Z[k]);
                    const double au = (Rm - Rk) / (Rm + Rk);
                    if (c->grid_size_adjust)
                        amat[(size_t)m * na + k] = std::min(0.5, std::max(-0.5, au / (au * au - 1.0)));
                }
            }
    }
    // functional-term launches: the k_rho block lists, 16 points per workgroup
    const int RS = c->rec_stride;
    std::vector<LaunchRange> gl;
    for (const LaunchRange& r : c->rho_launches) {
        LaunchRange g = r;
        g.lds = 0;
        if (g.cls == 4) g.cls = 1;  // the 64-point rho class of ldn <= 256 blocks: the gradient's ldn <= 256 kernel
        std::vector<int> ids(r.count);
        PDFXC_HIP(hipMemcpy(ids.data(), c->d_rho_work + r.offset, sizeof(int) * r.count, hipMemcpyDeviceToHost));
        for (int b : ids) {
            const BlockDesc& B = c->blocks[b];
            const size_t lds = sizeof(double) * (2 * (size_t)B.ldn * 16 + (size_t)B.nshell * RS + 3 * (size_t)na) +
                               sizeof(int) * (size_t)(B.ldn + 3 * B.nshell);
            if (g.cls > 2 || lds > (size_t)kMaxDynLds) {
                set_error("pdfxc_gradient_rks: a block with " + std::to_string(B.nao) + " AOs and " +
                          std::to_string(na) + " atoms does not fit the gradient kernel (raise ao_screen)");
                return PDFXC_ERR_UNSUPPORTED;
            }
            g.lds = std::max(g.lds, lds);
        }
        gl.push_back(g);
    }
    int *d_shat = nullptr, *d_slat = nullptr;
    double *d_at = nullptr, *d_ewf = nullptr, *d_g = nullptr;
    int rc = dalloc(&d_shat, shell_atom.size());
    if (rc == PDFXC_OK) rc = dalloc(&d_slat, slot_atom.size());
    if (rc == PDFXC_OK) rc = dalloc(&d_at, at.size());
    if (rc == PDFXC_OK) rc = dalloc(&d_ewf, (size_t)c->n_slots);
    if (rc == PDFXC_OK) rc = dalloc(&d_g, 6 * (size_t)na);
    hipError_t e = hipSuccess;
    if (rc == PDFXC_OK) {
        e = hipMemcpy(d_shat, shell_atom.data(), sizeof(int) * shell_atom.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_slat, slot_atom.data(), sizeof(int) * slot_atom.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_at, at.data(), sizeof(double) * at.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(d_g, 0, sizeof(double) * 6 * (size_t)na);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_gradient_rks: upload");
    }
    if (rc == PDFXC_OK) {
        double* H = c->h_pinned;
        for (int r = 0, k = 0; r < N; ++r)
            for (int q = 0; q <= r; ++q, ++k) {
                H[(size_t)r * N + q] = P_packed[k];
                H[(size_t)q * N + r] = P_packed[k];
            }
        e = hipMemcpyAsync(c->d_P, H, sizeof(double) * (size_t)N * N, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_gradient_rks: P upload");
    }
    // rho (the build's rho pass: FP64, or split when the context's precision
    // splits it), the point terms (with w e_xc per slot), the two gradient
    // terms (the functional term's contractions split in every split mode), E_xc
    c->rho_accumulate = 0;
    if (rc == PDFXC_OK) rc = run_rho(c, c->d_P, c->d_rho);
    if (rc == PDFXC_OK) rc = launch_xc_points(*c, functional, d_ewf);
    for (const LaunchRange& r : gl)
        if (rc == PDFXC_OK) rc = launch_grad_fn(*c, r, d_shat, d_slat, na, d_g);
    if (rc == PDFXC_OK) rc = launch_grad_w(*c, d_at, na, d_slat, d_ewf, d_g, c->grid_partition);
    if (rc == PDFXC_OK) rc = launch_energy_sum(*c, c->d_E);
    std::vector<double> g(6 * (size_t)na);
    double E = 0.0;
    if (rc == PDFXC_OK) {
        e = hipMemcpyAsync(g.data(), d_g, sizeof(double) * g.size(), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&E, c->d_E, sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_gradient_rks: results");
    }
    dfree(d_shat);
    dfree(d_slat);
    dfree(d_at);
    dfree(d_ewf);
    dfree(d_g);
    if (rc != PDFXC_OK) return rc;
    std::copy(g.begin(), g.begin() + 3 * na, grad_functional);
    std::copy(g.begin() + 3 * na, g.end(), grad_weight);
    if (E_xc) *E_xc = E;
    return PDFXC_OK;
}

int pdfxc_set_tuning(pdfxc_ctx* ctx, const pdfxc_tuning* t) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c || !t) return arg_error("pdfxc_set_tuning: bad arguments");
    if (t->rho_shift < 0 || t->rho_shift > 3 || t->split_rho_shift < 0 || t->split_rho_shift > 2 ||
        t->k_tiles < 8 || t->k_tiles > 96 || t->split_k_patches < 4 || t->split_k_patches > 32 ||
        (t->split_k_group != 1 && t->split_k_group != 2 && t->split_k_group != 4 && t->split_k_group != 8) ||
        t->persistent < 0 || t->persistent > 3)
        return arg_error("pdfxc_set_tuning: shape out of range");
    const pdfxc_tuning prev = c->tune;
    c->tune = *t;
    if (c->n_pts == 0) return PDFXC_OK;
    if (t->rho_shift == prev.rho_shift && t->k_tiles == prev.k_tiles && t->split_rho_shift == prev.split_rho_shift &&
        t->split_k_patches == prev.split_k_patches && t->split_k_group == prev.split_k_group)
        return PDFXC_OK;  // same blocks and work lists (the launch mode is chosen per launch)
    PDFXC_HIP(hipStreamSynchronize(c->stream));
    // a shape the grid's blocks do not fit drops the grid (build_blocks): put
    // the grid back with the previous shape and report the failure
    const std::vector<double> gx = c->gx, gy = c->gy, gz = c->gz, gw = c->gw;
    const std::vector<int> gen_atom = c->gen_atom;
    const int64_t n_pts = c->n_pts;
    const int rc = build_blocks(*c);
    if (rc == PDFXC_OK) return rc;
    const std::string msg = g_last_error;
    c->tune = prev;
    c->n_pts = n_pts;
    c->gx = gx;
    c->gy = gy;
    c->gz = gz;
    c->gw = gw;
    c->gen_atom = gen_atom;
    if (build_blocks(*c) != PDFXC_OK) c->tune = prev;
    set_error(msg);
    return rc;
}

int pdfxc_autotune(pdfxc_ctx* ctx, int functional, const double* dP, double* dK, double* dE, int reps,
                   pdfxc_tuning* chosen) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_autotune: null context");
    if (int rc = check_functional(functional, "pdfxc_autotune")) return rc;
    if (!dP || !dK || !dE || reps <= 0) return arg_error("pdfxc_autotune: bad arguments");
    if (c->n_pts == 0 || c->blocks.empty()) {
        set_error("pdfxc_autotune: no grid set");
        return PDFXC_ERR_NO_GRID;
    }
    std::vector<pdfxc_tuning> cand;
    if (c->opts.precision == PDFXC_FP64) {
        for (int rs : {0, 1})
            for (int kt : {96, 48, 24}) cand.push_back(pdfxc_tuning{rs, kt, 0, 24, 4, 2, 0.0});
    } else {
        // rho pass classes, then the split K pass's item size and group size
        // (points per K work unit) on the fastest rho shape
        for (int rs : {0, 1})
            cand.push_back(prec_split_rho(c->opts.precision) ? pdfxc_tuning{0, 96, rs, 24, 4, 2, 0.0}
                                                              : pdfxc_tuning{rs, 96, 0, 24, 4, 2, 0.0});
        for (int kp : {16, 24, 32})
            for (int kg : {1, 2, 4, 8})
                if (!(kp == 24 && kg == 4)) cand.push_back(pdfxc_tuning{-1, 96, -1, kp, kg, 2, 0.0});
    }
    // last: the fastest shape with the other launch modes (persistent rho
    // and / or K kernels, or one workgroup per item)
    for (int mask = 0; mask < 4; ++mask) cand.push_back(pdfxc_tuning{-2, -2, -2, -2, -2, mask, 0.0});
    pdfxc_tuning best = cand[0];
    best.ms = 1e30;
    for (pdfxc_tuning t : cand) {
        if (t.k_tiles == -2) {
            if (t.persistent == best.persistent) continue;  // timed already
            const int mask = t.persistent;
            t = best;
            t.persistent = mask;
        }
        if (t.rho_shift < 0) t.rho_shift = best.rho_shift;  // keep the best rho shape found so far
        if (t.split_rho_shift < 0) t.split_rho_shift = best.split_rho_shift;
        if (int rc = pdfxc_set_tuning(ctx, &t)) return rc;
        if (int rc = build_device(c, functional, dP, dK, dE)) return rc;  // warm-up
        double sum = 0.0;
        for (int r = 0; r < reps; ++r) {
            if (int rc = build_device(c, functional, dP, dK, dE)) return rc;
            sum += c->last_build_ms;
        }
        t.ms = sum / reps;
        if (t.ms < best.ms) best = t;
    }
    if (int rc = pdfxc_set_tuning(ctx, &best)) return rc;
    if (chosen) *chosen = best;
    return PDFXC_OK;
}

// Density (and its gradient) of P at arbitrary points: DfCalcGridX::gridDensity
// (DfCalcGridX.cpp:1038-1044) as TlDensField::makeDensFld calls it
// (src/tools/TlDensField.cpp:35-53).  The points go through the same block
// pipeline as a grid, in a context of their own, so the caller's grid and
// stored rho are untouched.
int pdfxc_density_at_points(pdfxc_ctx* ctx, const double* P_packed, int64_t n, const double* xyz, double* rho,
                            double* grad_rho) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_density_at_points: null context");
    if (n < 0 || (n > 0 && (!P_packed || !xyz || !rho))) return arg_error("pdfxc_density_at_points: bad arguments");
    if (n == 0) return PDFXC_OK;
    const int N = c->n_aos;
    std::vector<double> center(3 * (size_t)N), exps((size_t)N * kMaxPrim, 0.0), coefs((size_t)N * kMaxPrim, 0.0);
    std::vector<int> type(N), nprim(N);
    static const int first_type[3] = {0, 1, 4};
    for (const Shell& sh : c->shells)
        for (int m = 0; m < 2 * sh.l + 1; ++m) {
            const int a = sh.first_ao + m;
            type[a] = first_type[sh.l] + m;
            nprim[a] = sh.nprim;
            for (int d = 0; d < 3; ++d) center[3 * a + d] = sh.center[d];
            for (int k = 0; k < sh.nprim; ++k) {
                exps[(size_t)a * kMaxPrim + k] = sh.exps[k];
                coefs[(size_t)a * kMaxPrim + k] = sh.coefs[k];
            }
        }
    pdfxc_opts o = c->opts;
    o.precision = PDFXC_FP64;
    pdfxc_ctx* th = nullptr;
    int rc = pdfxc_ctx_create(&th, N, center.data(), type.data(), nprim.data(), exps.data(), coefs.data(), kMaxPrim,
                              &o);
    if (rc) return rc;
    Context* t = reinterpret_cast<Context*>(th);
    std::vector<double> X(n), Y(n), Zc(n), W(n, 1.0);
    for (int64_t i = 0; i < n; ++i) {
        X[i] = xyz[3 * i];
        Y[i] = xyz[3 * i + 1];
        Zc[i] = xyz[3 * i + 2];
    }
    rc = pdfxc_set_grid(th, n, X.data(), Y.data(), Zc.data(), W.data());
    if (rc == PDFXC_OK) {
        double* H = t->h_pinned;
        for (int r = 0, k = 0; r < N; ++r)
            for (int q = 0; q <= r; ++q, ++k) {
                H[(size_t)r * N + q] = P_packed[k];
                H[(size_t)q * N + r] = P_packed[k];
            }
        hipError_t e = hipMemcpyAsync(t->d_P, H, sizeof(double) * (size_t)N * N, hipMemcpyHostToDevice, t->stream);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_density_at_points: P upload");
    }
    if (rc == PDFXC_OK) rc = run_rho(t, t->d_P, t->d_rho);
    if (rc == PDFXC_OK) {
        std::vector<double> rs(4 * (size_t)t->n_slots);
        hipError_t e = hipStreamSynchronize(t->stream);
        if (e == hipSuccess) e = hipMemcpy(rs.data(), t->d_rho, rs.size() * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_density_at_points: download");
        for (int64_t i = 0; rc == PDFXC_OK && i < n; ++i) {
            const int64_t sl = t->slot_of_point[i];
            rho[i] = rs[4 * sl];
            if (grad_rho)
                for (int d = 0; d < 3; ++d) grad_rho[3 * i + d] = rs[4 * sl + 1 + d];
        }
    }
    pdfxc_ctx_destroy(th);
    return rc;
}

int pdfxc_eval_functional(int functional, int64_t n, const double* in, double* out) {
    if (int rc = check_functional(functional, "pdfxc_eval_functional")) return rc;
    if (n < 0 || (n > 0 && (!in || !out))) return arg_error("pdfxc_eval_functional: bad arguments");
    if (n == 0) return PDFXC_OK;
    if (functional == kHF) {
        std::fill(out, out + 4 * n, 0.0);
        return PDFXC_OK;
    }
    double *di = nullptr, *dout = nullptr;
    int rc = dalloc(&di, 2 * (size_t)n);
    if (rc == PDFXC_OK) rc = dalloc(&dout, 4 * (size_t)n);
    if (rc == PDFXC_OK) {
        hipError_t e = hipMemcpy(di, in, 2 * n * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_eval_functional: upload");
    }
    if (rc == PDFXC_OK) rc = launch_eval_functional(functional, n, di, dout, nullptr);
    if (rc == PDFXC_OK) {
        hipError_t e = hipMemcpy(out, dout, 4 * n * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_eval_functional: download");
    }
    dfree(di);
    dfree(dout);
    return rc;
}

int pdfxc_eval_functional_noalloc(int functional, int64_t n, const double* in, double* out) {
    if (int rc = check_functional(functional, "pdfxc_eval_functional_noalloc")) return rc;
    return eval_functional_noalloc(functional, n, in, out);
}

int pdfxc_eval_aos(pdfxc_ctx* ctx, int64_t n, const double* xyz, double* phi, double* gx, double* gy,
                   double* gz) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return arg_error("pdfxc_eval_aos: null context");
    if (n < 0 || (n > 0 && (!xyz || !phi))) return arg_error("pdfxc_eval_aos: bad arguments");
    if (n == 0) return PDFXC_OK;
    const size_t plane = (size_t)n * c->n_aos;
    double *dx = nullptr, *dout = nullptr;
    int rc = dalloc(&dx, 3 * (size_t)n);
    if (rc == PDFXC_OK) rc = dalloc(&dout, 4 * plane);
    if (rc == PDFXC_OK) {
        hipError_t e = hipMemcpy(dx, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_eval_aos: upload");
    }
    if (rc == PDFXC_OK) rc = launch_eval_aos_points(*c, n, dx, dout, c->stream);
    if (rc == PDFXC_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        double* outs[4] = {phi, gx, gy, gz};
        for (int q = 0; q < 4 && e == hipSuccess; ++q)
            if (outs[q]) e = hipMemcpy(outs[q], dout + q * plane, plane * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_eval_aos: download");
    }
    dfree(dx);
    dfree(dout);
    return rc;
}

int pdfxc_get_stats(pdfxc_ctx* ctx, pdfxc_stats* s) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c || !s) return arg_error("pdfxc_get_stats: bad arguments");
    s->n_points = c->n_pts;
    s->n_blocks = (int64_t)c->blocks.size();
    s->n_chunks = (c->opts.precision != PDFXC_FP64)
                      ? (int64_t)(c->srho_launches.size() + c->sk_launches.size() + c->k_fp64_launches.size() + 2)
                      : (int64_t)(c->rho_launches.size() + c->kmat_launches.size());
    s->flops = c->flops;
    s->mean_nao = c->mean_nao;
    s->max_nao = c->max_nao;
    s->last_build_ms = c->last_build_ms;
    for (int k = 0; k < 4; ++k) s->last_kernel_ms[k] = c->last_kernel_ms[k];
    s->n_k_fp64_blocks = c->n_k_fp64_blocks;
    return PDFXC_OK;
}

int pdfxc_selftest_mfma(double* out16x16) {
    if (!out16x16) return arg_error("pdfxc_selftest_mfma: null output");
    double* d = nullptr;
    int rc = dalloc(&d, 256);
    if (rc) return rc;
    rc = launch_mfma_selftest(d, nullptr);
    if (rc == PDFXC_OK) {
        hipError_t e = hipMemcpy(out16x16, d, 256 * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_selftest_mfma");
    }
    dfree(d);
    return rc;
}

int pdfxc_selftest_mfma_split(float* out4x16x16) {
    if (!out4x16x16) return arg_error("pdfxc_selftest_mfma_split: null output");
    float* d = nullptr;
    int rc = dalloc(&d, 1024);
    if (rc) return rc;
    rc = launch_mfma_split_selftest(d, nullptr);
    if (rc == PDFXC_OK) {
        hipError_t e = hipMemcpy(out4x16x16, d, 1024 * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_selftest_mfma_split");
    }
    dfree(d);
    return rc;
}

int pdfxc_selftest_copy(const double* d_src, double* d_dst, int64_t n) {
    if (n < 0 || (n > 0 && (!d_src || !d_dst))) return arg_error("pdfxc_selftest_copy: bad arguments");
    int rc = launch_copy(d_src, d_dst, n, nullptr);
    if (rc == PDFXC_OK) PDFXC_HIP(hipDeviceSynchronize());
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: several molecules' grid blocks in one set of launches.  Each
// molecule keeps its own AO indexing; its blocks carry the offset of its
// dense P and K in the concatenated buffers (BlockDesc::poff, ldp) and of its
// P slices (psoff), its slots start at a multiple of 256 so that the
// per-256-slot energy partials of one molecule are a contiguous range.  The
// launch sequence (P slicing per molecule, the rho classes, the point terms,
// the K classes, one energy sum per molecule) is captured once in a HIP graph
// and replayed.
// ---------------------------------------------------------------------------
namespace pdfxc {

struct Batch {
    Context m;                     // merged block lists, work lists and device buffers
    int n_mol = 0;
    std::vector<int> mol_n;        // AOs per molecule
    std::vector<int64_t> doff;     // dense P / K offsets (n_mol + 1)
    std::vector<int64_t> poff;     // packed offsets (n_mol + 1)
    std::vector<int64_t> mol_pts;  // points per molecule
    int* d_ebounds = nullptr;      // energy partial ranges (n_mol + 1)
    double* d_Pall = nullptr;      // host entry point: dense P and K of all molecules
    double* d_Kall = nullptr;
    double* d_Eall = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* key[3] = {nullptr, nullptr, nullptr};
    int key_xc = 0;
};

static void batch_free(Batch* b) {
    if (b->m.stream) (void)hipStreamSynchronize(b->m.stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    free_grid_buffers(&b->m);
    dfree(b->m.d_shrec);
    dfree(b->m.d_queue);
    dfree(b->m.d_Ps);
    dfree(b->m.d_perm);
    dfree(b->d_ebounds);
    dfree(b->d_Pall);
    dfree(b->d_Kall);
    dfree(b->d_Eall);
    for (auto& e : b->m.ev)
        if (e) (void)hipEventDestroy(e);
    if (b->m.stream) (void)hipStreamDestroy(b->m.stream);
    delete b;
}

// the build sequence on the batch's stream (no synchronisation: it is captured)
static int batch_sequence(Batch* b, int xc, const double* dP, double* dK, double* dE) {
    Context& c = b->m;
    QueueScope queues(&c);
    int rc = begin_queues(c);
    if (rc == PDFXC_OK) rc = run_rho(&c, dP, c.d_rho);
    if (rc == PDFXC_OK) rc = launch_xc_points(c, xc);
    if (rc == PDFXC_OK) rc = run_kmat(&c, c.d_coef, dK);
    if (rc == PDFXC_OK) rc = launch_energy_sums(c, b->d_ebounds, b->n_mol, dE);
    return rc;
}

}  // namespace pdfxc

extern "C" {

int pdfxc_batch_create(pdfxc_batch** out, int n_mol, pdfxc_ctx* const* ctxs) {
    if (!out || n_mol <= 0 || !ctxs) return arg_error("pdfxc_batch_create: bad arguments");
    *out = nullptr;
    for (int i = 0; i < n_mol; ++i) {
        const Context* c = reinterpret_cast<const Context*>(ctxs[i]);
        if (!c) return arg_error("pdfxc_batch_create: null context " + std::to_string(i));
        if (c->n_pts == 0 || c->blocks.empty()) {
            set_error("pdfxc_batch_create: context " + std::to_string(i) + " has no grid");
            return PDFXC_ERR_NO_GRID;
        }
        const Context* c0 = reinterpret_cast<const Context*>(ctxs[0]);
        if (c->opts.precision != c0->opts.precision || c->opts.density_cutoff != c0->opts.density_cutoff)
            return arg_error("pdfxc_batch_create: contexts differ in precision or density cut-off");
    }
    Batch* b = new Batch();
    Context& m = b->m;
    const Context* c0 = reinterpret_cast<const Context*>(ctxs[0]);
    m.opts = c0->opts;
    b->n_mol = n_mol;
    const bool split = m.opts.precision != PDFXC_FP64;
    int RS = 0;
    for (int i = 0; i < n_mol; ++i) RS = std::max(RS, reinterpret_cast<const Context*>(ctxs[i])->rec_stride);
    m.rec_stride = RS;
    std::vector<double> rec;
    std::vector<double> pts;
    std::vector<int> perm;
    std::vector<int> ebounds{0};
    std::vector<int> kw_cls;  // class of every FP64 KWork in m.h_k_work order (built per class below)
    b->doff.assign(1, 0);
    b->poff.assign(1, 0);
    int64_t psoff = 0;
    int shell_base = 0, nblocks = 0;
    int64_t slot_base = 0;
    std::vector<int> block_base(n_mol), slot_first(n_mol), group_base(n_mol, 0);
    for (int i = 0; i < n_mol; ++i) {
        const Context& c = *reinterpret_cast<const Context*>(ctxs[i]);
        const int N = c.n_aos;
        b->mol_n.push_back(N);
        b->mol_pts.push_back(c.n_pts);
        m.n_aos = std::max(m.n_aos, N);
        // shell records with the common stride
        for (const Shell& sh : c.shells) {
            const size_t o = rec.size();
            rec.resize(o + RS, 0.0);
            rec[o] = sh.center[0];
            rec[o + 1] = sh.center[1];
            rec[o + 2] = sh.center[2];
            rec[o + 3] = (double)(sh.l + 4 * sh.nprim);
            for (int q = 0; q < sh.nprim; ++q) {
                rec[o + 4 + 2 * q] = sh.exps[q];
                rec[o + 5 + 2 * q] = sh.coefs[q];
            }
        }
        // blocks, their lists and slots (the molecule's slots from a multiple of 256)
        block_base[i] = nblocks;
        slot_first[i] = (int)slot_base;
        const int sh_list = (int)m.block_shells.size(), col_list = (int)m.block_cols.size();
        const int chunk_list = (int)m.chunks.size();
        for (size_t k = 0; k < c.blocks.size(); ++k) {
            BlockDesc B = c.blocks[k];
            B.pt_begin += (int)slot_base;
            B.shell_begin += sh_list;
            B.col_begin += col_list;
            B.ldp = N;
            B.poff = b->doff.back();
            B.psoff = psoff;
            m.blocks.push_back(B);
            if (split) {
                SplitBlock SB = c.sblocks[k];
                SB.chunk_begin += chunk_list;
                m.sblocks.push_back(SB);
            }
        }
        for (int s : c.block_shells) m.block_shells.push_back(s + shell_base);
        m.block_shell_col.insert(m.block_shell_col.end(), c.block_shell_col.begin(), c.block_shell_col.end());
        m.block_sorder.insert(m.block_sorder.end(), c.block_sorder.begin(), c.block_sorder.end());
        m.block_cols.insert(m.block_cols.end(), c.block_cols.begin(), c.block_cols.end());
        if (split) {
            m.chunks.insert(m.chunks.end(), c.chunks.begin(), c.chunks.end());
            m.sshell_col.insert(m.sshell_col.end(), c.sshell_col.begin(), c.sshell_col.end());
            // split K groups with their lists moved to the batch's offsets
            group_base[i] = (int)m.kgroups.size();
            const int kb = (int)m.kg_blocks.size(), ks = (int)m.kg_shells.size(), kc = (int)m.kg_cols.size(),
                      km = (int)m.kg_masks.size();
            for (KGroup G : c.kgroups) {
                G.sub_begin += kb;
                G.shell_begin += ks;
                G.col_begin += kc;
                G.mask_begin += km;
                G.ldp = N;
                G.poff = b->doff.back();
                m.kgroups.push_back(G);
            }
            for (int x : c.kg_blocks) m.kg_blocks.push_back(x + block_base[i]);
            for (int x : c.kg_shells) m.kg_shells.push_back(x + shell_base);
            m.kg_shell_col.insert(m.kg_shell_col.end(), c.kg_shell_col.begin(), c.kg_shell_col.end());
            m.kg_sorder.insert(m.kg_sorder.end(), c.kg_sorder.begin(), c.kg_sorder.end());
            m.kg_cols.insert(m.kg_cols.end(), c.kg_cols.begin(), c.kg_cols.end());
            m.kg_masks.insert(m.kg_masks.end(), c.kg_masks.begin(), c.kg_masks.end());
            m.psplit_jobs.push_back(Context::PsplitJob{b->doff.back(), psoff, (int)perm.size(), N});
            perm.insert(perm.end(), c.split_perm.begin(), c.split_perm.end());
            psoff += psplit_elems(m.opts.precision, N);
        }
        pts.resize(4 * (size_t)(slot_base + c.n_slots), 0.0);
        for (int64_t p = 0; p < c.n_pts; ++p) {
            const int64_t sl = slot_base + c.slot_of_point[p];
            const int64_t ns = slot_base + c.n_slots;  // provisional column stride, fixed below
            (void)ns;
            pts[4 * sl + 0] = c.gx[p];
            pts[4 * sl + 1] = c.gy[p];
            pts[4 * sl + 2] = c.gz[p];
            pts[4 * sl + 3] = c.gw[p];
        }
        slot_base += c.n_slots;
        slot_base = (slot_base + 255) / 256 * 256;  // zero-weight padding slots
        ebounds.push_back((int)(slot_base / 256));
        nblocks += (int)c.blocks.size();
        shell_base += (int)c.shells.size();
        b->doff.push_back(b->doff.back() + (int64_t)N * N);
        b->poff.push_back(b->poff.back() + (int64_t)N * (N + 1) / 2);
        m.n_pts += c.n_pts;
        m.flops += c.flops;
        m.max_nao = std::max(m.max_nao, c.max_nao);
    }
    m.n_slots = slot_base;
    pts.resize(4 * (size_t)m.n_slots, 0.0);
    // SoA layout x | y | z | w as the kernels read it
    std::vector<double> soa(4 * (size_t)m.n_slots);
    for (int64_t sl = 0; sl < m.n_slots; ++sl)
        for (int q = 0; q < 4; ++q) soa[q * m.n_slots + sl] = pts[4 * sl + q];
    pts.clear();
    pts.shrink_to_fit();
    {
        double sum = 0.0;
        for (int i = 0; i < n_mol; ++i) {
            const Context& c = *reinterpret_cast<const Context*>(ctxs[i]);
            sum += c.mean_nao * (double)c.blocks.size();
        }
        m.mean_nao = sum / (double)nblocks;
    }
    // work lists: every launch class over all molecules
    auto merge_int = [&](auto get_ranges, auto get_list, std::vector<int>& out_list, std::vector<LaunchRange>& out_r) {
        for (int cls = 0; cls < kRhoFp64Class + 8; ++cls) {
            LaunchRange r{cls, (int)out_list.size(), 0, 0};
            for (int i = 0; i < n_mol; ++i) {
                const Context& c = *reinterpret_cast<const Context*>(ctxs[i]);
                for (const LaunchRange& q : get_ranges(c)) {
                    if (q.cls != cls) continue;
                    const std::vector<int>& L = get_list(c);
                    for (int k = 0; k < q.count; ++k) out_list.push_back(L[q.offset + k] + block_base[i]);
                    r.lds = std::max(r.lds, q.lds);
                }
            }
            r.count = (int)out_list.size() - r.offset;
            if (r.count) out_r.push_back(r);
        }
    };
    auto merge_kw = [&](auto get_ranges, auto get_list, std::vector<KWork>& out_list, std::vector<LaunchRange>& out_r) {
        for (int cls = 0; cls < 8; ++cls) {
            LaunchRange r{cls, (int)out_list.size(), 0, 0};
            for (int i = 0; i < n_mol; ++i) {
                const Context& c = *reinterpret_cast<const Context*>(ctxs[i]);
                for (const LaunchRange& q : get_ranges(c)) {
                    if (q.cls != cls) continue;
                    const std::vector<KWork>& L = get_list(c);
                    for (int k = 0; k < q.count; ++k) {
                        KWork w = L[q.offset + k];
                        w.block += block_base[i];
                        out_list.push_back(w);
                    }
                    r.lds = std::max(r.lds, q.lds);
                }
            }
            r.count = (int)out_list.size() - r.offset;
            if (r.count) out_r.push_back(r);
        }
    };
    merge_int([](const Context& c) -> const std::vector<LaunchRange>& { return c.rho_launches; },
              [](const Context& c) -> const std::vector<int>& { return c.h_rho_work; }, m.h_rho_work, m.rho_launches);
    merge_kw([](const Context& c) -> const std::vector<LaunchRange>& { return c.kmat_launches; },
             [](const Context& c) -> const std::vector<KWork>& { return c.h_k_work; }, m.h_k_work, m.kmat_launches);
    if (split) {
        merge_kw([](const Context& c) -> const std::vector<LaunchRange>& { return c.k_fp64_launches; },
                 [](const Context& c) -> const std::vector<KWork>& { return c.h_k_work; }, m.h_k_work,
                 m.k_fp64_launches);
        merge_int([](const Context& c) -> const std::vector<LaunchRange>& { return c.srho_launches; },
                  [](const Context& c) -> const std::vector<int>& { return c.h_srho_work; }, m.h_srho_work,
                  m.srho_launches);
        for (int cls = 0; cls < 8; ++cls) {  // split K items, every molecule's of one shape together
            LaunchRange r{cls, (int)m.h_kg_work.size(), 0, 0};
            for (int i = 0; i < n_mol; ++i) {
                const Context& c = *reinterpret_cast<const Context*>(ctxs[i]);
                for (const LaunchRange& q : c.sk_launches) {
                    if (q.cls != cls) continue;
                    for (int k = 0; k < q.count; ++k) {
                        KGWork w = c.h_kg_work[q.offset + k];
                        w.group += group_base[i];
                        m.h_kg_work.push_back(w);
                    }
                    r.lds = std::max(r.lds, q.lds);
                }
            }
            r.count = (int)m.h_kg_work.size() - r.offset;
            if (r.count) m.sk_launches.push_back(r);
        }
        for (int i = 0; i < n_mol; ++i) m.n_k_fp64_blocks += reinterpret_cast<const Context*>(ctxs[i])->n_k_fp64_blocks;
    }
    (void)kw_cls;
    (void)slot_first;
    // device side
    hipError_t e = hipGetDevice(&m.device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&m.n_cu, hipDeviceAttributeMultiprocessorCount, m.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking);
    for (int k = 0; k < 6 && e == hipSuccess; ++k) e = hipEventCreate(&m.ev[k]);
    int rc = (e == hipSuccess) ? PDFXC_OK : hip_fail(e, "pdfxc_batch_create: stream/event creation");
    auto up = [&](auto* d, const auto& v) {
        if (rc != PDFXC_OK || v.empty()) return;
        const hipError_t ee = hipMemcpy(d, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice);
        if (ee != hipSuccess) rc = hip_fail(ee, "pdfxc_batch_create: upload");
    };
    if (rc == PDFXC_OK) rc = dalloc(&m.d_shrec, rec.size());
    if (rc == PDFXC_OK) rc = dalloc(&m.d_queue, kMaxQueues);
    if (rc == PDFXC_OK) rc = dalloc(&m.d_pts, soa.size());
    if (rc == PDFXC_OK) rc = dalloc(&m.d_blocks, m.blocks.size());
    if (rc == PDFXC_OK) rc = dalloc(&m.d_block_shells, std::max<size_t>(1, m.block_shells.size()));
    if (rc == PDFXC_OK) rc = dalloc(&m.d_block_cols, std::max<size_t>(1, m.block_cols.size()));
    if (rc == PDFXC_OK) rc = dalloc(&m.d_block_shell_col, std::max<size_t>(1, m.block_shell_col.size()));
    if (rc == PDFXC_OK) rc = dalloc(&m.d_block_sorder, std::max<size_t>(1, m.block_sorder.size()));
    if (rc == PDFXC_OK) rc = dalloc(&m.d_rho_work, std::max<size_t>(1, m.h_rho_work.size()));
    if (rc == PDFXC_OK) rc = dalloc(reinterpret_cast<char**>(&m.d_k_work), std::max<size_t>(1, m.h_k_work.size()) * sizeof(KWork));
    if (rc == PDFXC_OK) rc = dalloc(&m.d_coef, 4 * (size_t)m.n_slots);
    if (rc == PDFXC_OK) rc = dalloc(&m.d_rho, 4 * (size_t)m.n_slots);
    if (rc == PDFXC_OK) rc = dalloc(&m.d_eblock, (size_t)m.n_slots / 256);
    if (rc == PDFXC_OK) rc = dalloc(&b->d_ebounds, ebounds.size());
    if (rc == PDFXC_OK && split) {
        rc = dalloc(&m.d_sblocks, m.sblocks.size());
        if (rc == PDFXC_OK) rc = dalloc(&m.d_chunks, std::max<size_t>(1, m.chunks.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_sshell_col, std::max<size_t>(1, m.sshell_col.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_srho_work, std::max<size_t>(1, m.h_srho_work.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_kg_work, std::max<size_t>(1, m.h_kg_work.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_kgroups, std::max<size_t>(1, m.kgroups.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_kg_blocks, std::max<size_t>(1, m.kg_blocks.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_kg_shells, std::max<size_t>(1, m.kg_shells.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_kg_shell_col, std::max<size_t>(1, m.kg_shell_col.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_kg_sorder, std::max<size_t>(1, m.kg_sorder.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_kg_cols, std::max<size_t>(1, m.kg_cols.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_kg_masks, std::max<size_t>(1, m.kg_masks.size()));
        if (rc == PDFXC_OK) rc = dalloc(&m.d_perm, perm.size());
        if (rc == PDFXC_OK) rc = dalloc(&m.d_Ps, (size_t)psoff);
    }
    up(m.d_shrec, rec);
    up(m.d_pts, soa);
    up(m.d_blocks, m.blocks);
    up(m.d_block_shells, m.block_shells);
    up(m.d_block_cols, m.block_cols);
    up(m.d_block_shell_col, m.block_shell_col);
    up(m.d_block_sorder, m.block_sorder);
    up(m.d_rho_work, m.h_rho_work);
    up(reinterpret_cast<KWork*>(m.d_k_work), m.h_k_work);
    up(b->d_ebounds, ebounds);
    if (split) {
        up(m.d_sblocks, m.sblocks);
        up(m.d_chunks, m.chunks);
        up(m.d_sshell_col, m.sshell_col);
        up(m.d_srho_work, m.h_srho_work);
        up(m.d_kg_work, m.h_kg_work);
        up(m.d_kgroups, m.kgroups);
        up(m.d_kg_blocks, m.kg_blocks);
        up(m.d_kg_shells, m.kg_shells);
        up(m.d_kg_shell_col, m.kg_shell_col);
        up(m.d_kg_sorder, m.kg_sorder);
        up(m.d_kg_cols, m.kg_cols);
        up(m.d_kg_masks, m.kg_masks);
        up(m.d_perm, perm);
    }
    if (rc == PDFXC_OK) {
        e = hipMemset(m.d_rho, 0, 4 * (size_t)m.n_slots * sizeof(double));
        if (e == hipSuccess) e = hipMemset(m.d_coef, 0, 4 * (size_t)m.n_slots * sizeof(double));
        if (e != hipSuccess) rc = hip_fail(e, "pdfxc_batch_create: clear");
    }
    if (rc != PDFXC_OK) {
        batch_free(b);
        return rc;
    }
    *out = reinterpret_cast<pdfxc_batch*>(b);
    return PDFXC_OK;
}

int pdfxc_batch_destroy(pdfxc_batch* batch) {
    if (batch) batch_free(reinterpret_cast<Batch*>(batch));
    return PDFXC_OK;
}

int pdfxc_batch_offsets(pdfxc_batch* batch, int64_t* dense_off, int64_t* packed_off) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return arg_error("pdfxc_batch_offsets: null batch");
    if (dense_off) std::copy(b->doff.begin(), b->doff.end(), dense_off);
    if (packed_off) std::copy(b->poff.begin(), b->poff.end(), packed_off);
    return PDFXC_OK;
}

int pdfxc_batch_build_rks_device(pdfxc_batch* batch, int functional, const double* dP, double* dK, double* dE) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return arg_error("pdfxc_batch_build_rks_device: null batch");
    if (int rc = check_functional(functional, "pdfxc_batch_build_rks_device")) return rc;
    if (!dP || !dK || !dE) return arg_error("pdfxc_batch_build_rks_device: null buffer");
    Context& c = b->m;
    if (functional == kHF) {
        if (int rc = launch_zero(dE, b->n_mol, c.stream)) return rc;
        PDFXC_HIP(hipStreamSynchronize(c.stream));
        return PDFXC_OK;
    }
    if (!b->exec || b->key[0] != dP || b->key[1] != dK || b->key[2] != dE || b->key_xc != functional) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        PDFXC_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_sequence(b, functional, dP, dK, dE);
        const hipError_t ec = hipStreamEndCapture(c.stream, &b->graph);
        if (rc) return rc;
        if (ec != hipSuccess) return hip_fail(ec, "pdfxc_batch_build_rks_device: capture");
        PDFXC_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->key[0] = dP;
        b->key[1] = dK;
        b->key[2] = dE;
        b->key_xc = functional;
    }
    PDFXC_HIP(hipEventRecord(c.ev[0], c.stream));
    PDFXC_HIP(hipGraphLaunch(b->exec, c.stream));
    PDFXC_HIP(hipEventRecord(c.ev[1], c.stream));
    PDFXC_HIP(hipEventSynchronize(c.ev[1]));
    float t = 0.f;
    PDFXC_HIP(hipEventElapsedTime(&t, c.ev[0], c.ev[1]));
    c.last_build_ms = t;
    return PDFXC_OK;
}

int pdfxc_batch_build_rks(pdfxc_batch* batch, int functional, const double* P_packed, double* K_packed, double* E) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return arg_error("pdfxc_batch_build_rks: null batch");
    if (int rc = check_functional(functional, "pdfxc_batch_build_rks")) return rc;
    if (!P_packed || !K_packed || !E) return arg_error("pdfxc_batch_build_rks: null buffer");
    Context& c = b->m;
    const int64_t total = b->doff.back();
    if (!b->d_Pall) {
        if (int rc = dalloc(&b->d_Pall, (size_t)total)) return rc;
        if (int rc = dalloc(&b->d_Kall, (size_t)total)) return rc;
        if (int rc = dalloc(&b->d_Eall, (size_t)b->n_mol)) return rc;
    }
    std::vector<double> H;
    for (int i = 0; i < b->n_mol; ++i) {
        const int N = b->mol_n[i];
        H.assign((size_t)N * N, 0.0);
        const double* pp = P_packed + b->poff[i];
        for (int r = 0, k = 0; r < N; ++r)
            for (int q = 0; q <= r; ++q, ++k) H[(size_t)r * N + q] = H[(size_t)q * N + r] = pp[k];
        PDFXC_HIP(hipMemcpy(b->d_Pall + b->doff[i], H.data(), H.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    PDFXC_HIP(hipMemsetAsync(b->d_Kall, 0, (size_t)total * sizeof(double), c.stream));
    PDFXC_HIP(hipStreamSynchronize(c.stream));
    if (int rc = pdfxc_batch_build_rks_device(batch, functional, b->d_Pall, b->d_Kall, b->d_Eall)) return rc;
    if (functional == kHF) {
        std::fill(E, E + b->n_mol, 0.0);
        return PDFXC_OK;
    }
    PDFXC_HIP(hipMemcpy(E, b->d_Eall, sizeof(double) * b->n_mol, hipMemcpyDeviceToHost));
    for (int i = 0; i < b->n_mol; ++i) {
        const int N = b->mol_n[i];
        H.resize((size_t)N * N);
        PDFXC_HIP(hipMemcpy(H.data(), b->d_Kall + b->doff[i], H.size() * sizeof(double), hipMemcpyDeviceToHost));
        double* kp = K_packed + b->poff[i];
        for (int r = 0, k = 0; r < N; ++r)
            for (int q = 0; q <= r; ++q, ++k) kp[k] += H[(size_t)r * N + q];
    }
    return PDFXC_OK;
}

int pdfxc_batch_stats(pdfxc_batch* batch, pdfxc_stats* s) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b || !s) return arg_error("pdfxc_batch_stats: bad arguments");
    const Context& c = b->m;
    std::memset(s, 0, sizeof(*s));
    s->n_points = c.n_pts;
    s->n_blocks = (int64_t)c.blocks.size();
    s->n_chunks = (int64_t)((c.opts.precision != PDFXC_FP64)
                                ? c.psplit_jobs.size() + c.srho_launches.size() + c.sk_launches.size() +
                                      c.k_fp64_launches.size() + 2
                                : c.rho_launches.size() + c.kmat_launches.size() + 2);
    s->flops = c.flops;
    s->mean_nao = c.mean_nao;
    s->max_nao = c.max_nao;
    s->last_build_ms = c.last_build_ms;
    s->n_k_fp64_blocks = c.n_k_fp64_blocks;
    return PDFXC_OK;
}

}  // extern "C"
