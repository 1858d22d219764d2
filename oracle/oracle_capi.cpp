// oracle_capi.cpp -- extern "C" surface of the parity oracle (partia_oracle.hpp).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py, as the checker.  Never part of the product.
//
// A handle holds one InterpolationInput after LinearBVH! (reference
// KI/interpolation_setup/InterpolationInput.jl:201-221): particle columns in
// Morton order, the Morton encoding and the tree.  precision 64 runs the
// reference algorithm with T = Float64, precision 32 with T = Float32 (inputs
// converted first, as build_input would for Float32 columns).  The sample loops
// are OpenMP `parallel for schedule(dynamic)`, the counterpart of the
// reference's `@threads` loops (KI/grid_interpolation/drivers/PointSamples_driver.jl:38,
// LineSamples_driver.jl:50).
#include <omp.h>

#include <cstring>
#include <memory>
#include <string>

#include "partia_oracle.hpp"
#include "oracle_dual.hpp"
#include "../partia.jl_amd/csrc/line_lut_tables.inc"

using namespace oracle;

namespace {

struct Handle {
    int precision = 64;
    int code_bits = 63;
    Input<double> d;
    Input<float> f;
    std::unique_ptr<LineLUT<double>> lut_d;
    std::unique_ptr<LineLUT<float>> lut_f;
};

thread_local std::string g_err;

Catalog make_catalog(int nS, const int* sslots, const unsigned char* sflags, int nG, const int* gslots, int nD,
                     const int* dslots, int nC, const int* cslots) {
    Catalog c;
    c.sslots.assign(sslots, sslots + nS);
    c.sflags.assign(sflags, sflags + nS);
    c.gslots.assign(gslots, gslots + nG);
    c.dslots.assign(dslots, dslots + 3 * nD);
    c.cslots.assign(cslots, cslots + 3 * nC);
    return c;
}

template <class T> void fill_input(Input<T>& in, int64_t n, const double* x, const double* y, const double* z,
                                   const double* m, const double* h, const double* rho, int ncol,
                                   const double* const* cols, int kernel) {
    auto cp = [&](const double* s) { return std::vector<T>(s, s + n); };
    in.kernel = kernel;
    in.x = std::vector<T>(x, x + n);
    in.y = cp(y); in.z = cp(z); in.m = cp(m); in.h = cp(h); in.rho = cp(rho);
    in.quant.clear();
    for (int c = 0; c < ncol; ++c) in.quant.push_back(cp(cols[c]));
}

// The float64 input with a derivative direction: tan holds (6 + ncol) rows of n
// derivatives in input order (x, y, z, m, h, rho, the quantity columns).  The
// tree keeps its topology; its leaves carry the coordinate and h derivatives,
// its node boxes none (they only steer the walk).
Input<Dual> dual_input(const Input<double>& in, const double* tan) {
    const int64_t n = (int64_t)in.x.size();
    auto col = [&](const std::vector<double>& v, int row) {
        std::vector<Dual> w((size_t)n);
        for (int64_t i = 0; i < n; ++i) w[(size_t)i] = Dual{v[(size_t)i], tan[(int64_t)row * n + (int64_t)in.order[(size_t)i] - 1]};
        return w;
    };
    auto val = [](const std::vector<double>& v) { return std::vector<Dual>(v.begin(), v.end()); };
    Input<Dual> o;
    o.kernel = in.kernel;
    o.dims = in.dims;
    o.order = in.order;
    o.codes = in.codes;
    o.x = col(in.x, 0);
    o.y = col(in.y, 1);
    o.z = col(in.z, 2);
    o.m = col(in.m, 3);
    o.h = col(in.h, 4);
    o.rho = col(in.rho, 5);
    for (size_t c = 0; c < in.quant.size(); ++c) o.quant.push_back(col(in.quant[c], 6 + (int)c));
    o.bvh.brt = in.bvh.brt;
    for (int d = 0; d < 3; ++d) {
        o.bvh.leaf[d] = col(in.bvh.leaf[d], d);
        o.bvh.nmin[d] = val(in.bvh.nmin[d]);
        o.bvh.nmax[d] = val(in.bvh.nmax[d]);
    }
    o.bvh.leaf_scale = col(in.bvh.leaf_scale, 4);
    o.bvh.nscale = val(in.bvh.nscale);
    return o;
}

template <class T> const LBVH<T>& bvh_of(const Handle* h);
template <> const LBVH<double>& bvh_of<double>(const Handle* h) { return h->d.bvh; }
template <> const LBVH<float>& bvh_of<float>(const Handle* h) { return h->f.bvh; }

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

void* oracle_build(int precision, int64_t n, const double* x, const double* y, const double* z, const double* m,
                   const double* h, const double* rho, int ncol, const double* const* cols, int kernel,
                   int code_bits) {
    try {
        auto H = std::make_unique<Handle>();
        H->precision = precision;
        H->code_bits = code_bits;
        if (precision == 64) {
            fill_input(H->d, n, x, y, z, m, h, rho, ncol, cols, kernel);
            if (code_bits == 30) linear_bvh_inplace<double, u32>(H->d);
            else linear_bvh_inplace<double, u64>(H->d);
            H->lut_d.reset(new LineLUT<double>(ph_lut_inorm));
        } else {
            fill_input(H->f, n, x, y, z, m, h, rho, ncol, cols, kernel);
            if (code_bits == 30) linear_bvh_inplace<float, u32>(H->f);
            else linear_bvh_inplace<float, u64>(H->f);
            H->lut_f.reset(new LineLUT<float>(ph_lut_inorm));
        }
        return H.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// LinearBVH! of an InterpolationInput{2}: as oracle_build with z all zeros
// and 2D Morton codes.
void* oracle_build2d(int precision, int64_t n, const double* x, const double* y, const double* m, const double* h,
                     const double* rho, int ncol, const double* const* cols, int kernel, int code_bits) {
    try {
        auto H = std::make_unique<Handle>();
        H->precision = precision;
        H->code_bits = code_bits;
        const std::vector<double> z((size_t)std::max<int64_t>(n, 0), 0.0);
        if (precision == 64) {
            fill_input(H->d, n, x, y, z.data(), m, h, rho, ncol, cols, kernel);
            H->d.dims = 2;
            if (code_bits == 30) linear_bvh_inplace<double, u32>(H->d);
            else linear_bvh_inplace<double, u64>(H->d);
            H->lut_d.reset(new LineLUT<double>(ph_lut_inorm));
        } else {
            fill_input(H->f, n, x, y, z.data(), m, h, rho, ncol, cols, kernel);
            H->f.dims = 2;
            if (code_bits == 30) linear_bvh_inplace<float, u32>(H->f);
            else linear_bvh_inplace<float, u64>(H->f);
            H->lut_f.reset(new LineLUT<float>(ph_lut_inorm));
        }
        return H.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void oracle_free(void* h) { delete static_cast<Handle*>(h); }

// Tree and encoding arrays (unified IDs, reference layout).  Arrays of length
// 2n-1 for left/right/escape/parent; n for order/codes.
int oracle_get_tree(void* hv, uint64_t* order, uint64_t* codes, int32_t* left, int32_t* right, int32_t* escape,
                    int32_t* parent, int32_t* root) {
    Handle* H = static_cast<Handle*>(hv);
    const std::vector<u64>& o = H->precision == 64 ? H->d.order : H->f.order;
    const std::vector<u64>& c = H->precision == 64 ? H->d.codes : H->f.codes;
    const BRT& t = H->precision == 64 ? H->d.bvh.brt : H->f.bvh.brt;
    std::memcpy(order, o.data(), o.size() * 8);
    std::memcpy(codes, c.data(), c.size() * 8);
    std::memcpy(left, t.left.data(), t.left.size() * 4);
    std::memcpy(right, t.right.data(), t.right.size() * 4);
    std::memcpy(escape, t.escape.data(), t.escape.size() * 4);
    std::memcpy(parent, t.parent.data(), t.parent.size() * 4);
    *root = t.root;
    return 0;
}

// Leaf coordinates (3 x n), leaf scale (n), node AABB (3 x (n-1) each), node
// scale (n-1), returned as float64 whatever the working precision.
int oracle_get_bvh(void* hv, double* leaf3, double* leaf_scale, double* nmin3, double* nmax3, double* nscale) {
    Handle* H = static_cast<Handle*>(hv);
    auto put = [&](const auto& b) {
        size_t n = b.leaf_scale.size(), ni = b.nscale.size();
        for (int d = 0; d < 3; ++d)
            for (size_t i = 0; i < n; ++i) leaf3[d * n + i] = double(b.leaf[d][i]);
        for (size_t i = 0; i < n; ++i) leaf_scale[i] = double(b.leaf_scale[i]);
        for (int d = 0; d < 3; ++d)
            for (size_t i = 0; i < ni; ++i) {
                nmin3[d * ni + i] = double(b.nmin[d][i]);
                nmax3[d * ni + i] = double(b.nmax[d][i]);
            }
        for (size_t i = 0; i < ni; ++i) nscale[i] = double(b.nscale[i]);
    };
    if (H->precision == 64) put(H->d.bvh);
    else put(H->f.bvh);
    return 0;
}

// Permuted particle columns (Morton order): x, y, z, m, h, rho, then ncol columns.
int oracle_get_columns(void* hv, double* out /* (6+ncol) x n */) {
    Handle* H = static_cast<Handle*>(hv);
    auto put = [&](const auto& in) {
        size_t n = in.x.size();
        using V = std::decay_t<decltype(in.x)>;
        const V* cols6[6] = {&in.x, &in.y, &in.z, &in.m, &in.h, &in.rho};
        for (int c = 0; c < 6; ++c)
            for (size_t i = 0; i < n; ++i) out[c * n + i] = double((*cols6[c])[i]);
        for (size_t c = 0; c < in.quant.size(); ++c)
            for (size_t i = 0; i < n; ++i) out[(6 + c) * n + i] = double(in.quant[c][i]);
    };
    if (H->precision == 64) put(H->d);
    else put(H->f);
    return 0;
}

// PointSamples_interpolation's per-sample loop: NN-h (unless scatter) then the
// fused kernel; out is L x S row-major.
int oracle_point_samples(void* hv, int nS, const int* sslots, const unsigned char* sflags, int nG, const int* gslots,
                         int nD, const int* dslots, int nC, const int* cslots, int strategy, int64_t S,
                         const double* px, const double* py, const double* pz, double* out, int nthreads) {
    Handle* H = static_cast<Handle*>(hv);
    Catalog cat = make_catalog(nS, sslots, sflags, nG, gslots, nD, dslots, nC, cslots);
    const int L = cat.L();
    auto run = [&](auto& in) {
        using T = std::decay_t<decltype(in.x[0])>;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        {
            std::vector<T> tmp(L);
#pragma omp for schedule(dynamic, 64)
            for (int64_t s = 0; s < S; ++s) {
                T p[3] = {T(px[s]), T(py[s]), T(pz[s])};
                T ha = strategy == SCATTER ? T(0) : nearest_h<T>(in.bvh, p);
                general_interpolate<T>(in, cat, strategy, p, ha, tmp.data(), 1);
                for (int k = 0; k < L; ++k) out[(int64_t)k * S + s] = double(tmp[k]);
            }
        }
    };
    if (H->precision == 64) run(H->d);
    else run(H->f);
    return 0;
}

// LineSamples_interpolation's per-ray loop; ha (per ray, may be null) is used
// only by the gather / symmetric kernels, which the drivers never call.
int oracle_line_samples(void* hv, int nS, const int* sslots, const unsigned char* sflags, int strategy, int64_t S,
                        const double* ox, const double* oy, const double* oz, const double* mx, const double* my,
                        const double* mz, const double* ha, double* out, int nthreads) {
    Handle* H = static_cast<Handle*>(hv);
    Catalog cat = make_catalog(nS, sslots, sflags, 0, nullptr, 0, nullptr, 0, nullptr);
    auto run = [&](auto& in, const auto& lut) {
        using T = std::decay_t<decltype(in.x[0])>;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        {
            std::vector<T> tmp(nS > 0 ? nS : 1);
#pragma omp for schedule(dynamic, 64)
            for (int64_t s = 0; s < S; ++s) {
                T o[3] = {T(ox[s]), T(oy[s]), T(oz[s])};
                T m[3] = {T(mx[s]), T(my[s]), T(mz[s])};
                T h = ha ? T(ha[s]) : T(0);
                line_interpolate<T>(in, lut, cat, strategy, o, m, h, tmp.data(), 1);
                for (int k = 0; k < nS; ++k) out[(int64_t)k * S + s] = double(tmp[k]);
            }
        }
    };
    if (H->precision == 64) run(H->d, *H->lut_d);
    else run(H->f, *H->lut_f);
    return 0;
}

// The single-quantity point kernels (single_interpolate): out is 1 x S (kind 0
// density, 1 number density) or 3 x S (kind 2 density gradient).  ha (per
// point, may be null) is the caller's smoothing length for gather / symmetric;
// null takes each point's nearest-particle h, as the drivers do.
int oracle_point_single(void* hv, int kind, int strategy, int64_t S, const double* px, const double* py,
                        const double* pz, const double* ha, double* out, int nthreads) {
    Handle* H = static_cast<Handle*>(hv);
    const int L = kind == 2 ? 3 : 1;
    auto run = [&](auto& in) {
        using T = std::decay_t<decltype(in.x[0])>;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        {
            T tmp[3];
#pragma omp for schedule(dynamic, 64)
            for (int64_t s = 0; s < S; ++s) {
                T p[3] = {T(px[s]), T(py[s]), T(pz[s])};
                T h = strategy == SCATTER ? T(0) : ha ? T(ha[s]) : nearest_h<T>(in.bvh, p);
                single_interpolate<T>(in, kind, strategy, p, h, tmp, 1);
                for (int k = 0; k < L; ++k) out[(int64_t)k * S + s] = double(tmp[k]);
            }
        }
    };
    if (H->precision == 64) run(H->d);
    else run(H->f);
    return 0;
}

// _line_integrated_density_kernel per line (line_density); ha as in
// oracle_line_samples.
int oracle_line_density(void* hv, int strategy, int64_t S, const double* ox, const double* oy, const double* oz,
                        const double* mx, const double* my, const double* mz, const double* ha, double* out,
                        int nthreads) {
    Handle* H = static_cast<Handle*>(hv);
    auto run = [&](auto& in, const auto& lut) {
        using T = std::decay_t<decltype(in.x[0])>;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : omp_get_max_threads()) schedule(dynamic, 64)
        for (int64_t s = 0; s < S; ++s) {
            T o[3] = {T(ox[s]), T(oy[s]), T(oz[s])};
            T m[3] = {T(mx[s]), T(my[s]), T(mz[s])};
            out[s] = double(line_density<T>(in, lut, strategy, o, m, ha ? T(ha[s]) : T(0)));
        }
    };
    if (H->precision == 64) run(H->d, *H->lut_d);
    else run(H->f, *H->lut_f);
    return 0;
}

// Adjoint of oracle_line_samples for catalog scalar `slot` (scatter): grad is
// 6 x n float64 in input order (d/dm, d/dA, d/dh, d/dx, d/dy, d/dz), summed over
// the rays with cotangents cot[S].  Serial (the rays share particles).
int oracle_line_adjoint(void* hv, int slot, int shepard, int64_t S, const double* ox, const double* oy,
                        const double* oz, const double* mx, const double* my, const double* mz, const double* cot,
                        double* grad) {
    Handle* H = static_cast<Handle*>(hv);
    auto run = [&](auto& in, const auto& lut) {
        using T = std::decay_t<decltype(in.x[0])>;
        const int64_t n = (int64_t)in.x.size();
        std::vector<T> g(6 * (size_t)n, T(0));
        for (int64_t s = 0; s < S; ++s) {
            T o[3] = {T(ox[s]), T(oy[s]), T(oz[s])};
            T m[3] = {T(mx[s]), T(my[s]), T(mz[s])};
            line_adjoint<T>(in, lut, slot, shepard != 0, o, m, T(cot[s]), g.data());
        }
        for (int c = 0; c < 6; ++c)
            for (int64_t i = 0; i < n; ++i) grad[c * n + (int64_t)in.order[i] - 1] = double(g[c * n + i]);
    };
    if (H->precision == 64) run(H->d, *H->lut_d);
    else run(H->f, *H->lut_f);
    return 0;
}

int oracle_point_adjoint(void* hv, int slot, int shepard, int strategy, int64_t S, const double* px,
                         const double* py, const double* pz, const double* cot, double* grad) {
    Handle* H = static_cast<Handle*>(hv);
    auto run = [&](auto& in) {
        using T = std::decay_t<decltype(in.x[0])>;
        const int64_t n = (int64_t)in.x.size();
        std::vector<T> g(6 * (size_t)n, T(0));
        for (int64_t s = 0; s < S; ++s) {
            T p[3] = {T(px[s]), T(py[s]), T(pz[s])};
            point_adjoint<T>(in, slot, shepard != 0, strategy, p, T(cot[s]), g.data());
        }
        for (int c = 0; c < 6; ++c)
            for (int64_t i = 0; i < n; ++i) grad[c * n + (int64_t)in.order[i] - 1] = double(g[c * n + i]);
    };
    if (H->precision == 64) run(H->d);
    else run(H->f);
    return 0;
}

// field_adjoint over S points; cot: S x (1 or 3) row-major; grad: (5 + c) x n, input order.
int oracle_field_adjoint(void* hv, int mode, const int* slots, int shepard, int strategy, int64_t S,
                         const double* px, const double* py, const double* pz, const double* cot, double* grad) {
    Handle* H = static_cast<Handle*>(hv);
    const int c = mode >= 2 ? 3 : 1, w = mode == 0 || mode == 2 ? 1 : 3, rows = 5 + c;
    auto run = [&](auto& in) {
        using T = std::decay_t<decltype(in.x[0])>;
        const int64_t n = (int64_t)in.x.size();
        std::vector<T> g((size_t)rows * (size_t)n, T(0));
        for (int64_t s = 0; s < S; ++s) {
            T p[3] = {T(px[s]), T(py[s]), T(pz[s])};
            T ct[3] = {T(cot[s * w]), T(w == 3 ? cot[s * w + 1] : 0), T(w == 3 ? cot[s * w + 2] : 0)};
            field_adjoint<T>(in, mode, slots, shepard != 0, strategy, p, ct, g.data());
        }
        for (int r = 0; r < rows; ++r)
            for (int64_t i = 0; i < n; ++i) grad[r * n + (int64_t)in.order[i] - 1] = double(g[r * n + i]);
    };
    if (H->precision == 64) run(H->d);
    else run(H->f);
    return 0;
}

double oracle_kernel_d2f(int K, double q) { return kernel_d2f<double>(K, q); }

// Forward mode through the float64 point path (nearest-particle h_a included):
// out (L x S) the outputs, dout (L x S) their derivatives along tan ((6 + ncol)
// x n, input order: x, y, z, m, h, rho, quantities).  Float64 handles only.
int oracle_point_jvp(void* hv, int nS, const int* sslots, const unsigned char* sflags, int nG, const int* gslots,
                     int nD, const int* dslots, int nC, const int* cslots, int strategy, int64_t S, const double* px,
                     const double* py, const double* pz, const double* tan, double* out, double* dout, int nthreads) {
    Handle* H = static_cast<Handle*>(hv);
    if (H->precision != 64) {
        g_err = "oracle_point_jvp needs a float64 handle";
        return 1;
    }
    Catalog cat = make_catalog(nS, sslots, sflags, nG, gslots, nD, dslots, nC, cslots);
    const int L = cat.L();
    const Input<Dual> in = dual_input(H->d, tan);
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    {
        std::vector<Dual> tmp((size_t)L);
#pragma omp for schedule(dynamic, 16)
        for (int64_t s = 0; s < S; ++s) {
            Dual p[3] = {Dual(px[s]), Dual(py[s]), Dual(pz[s])};
            Dual ha = strategy == SCATTER ? Dual(0.0) : nearest_h<Dual>(in.bvh, p);
            general_interpolate<Dual>(in, cat, strategy, p, ha, tmp.data(), 1);
            for (int k = 0; k < L; ++k) {
                out[(int64_t)k * S + s] = tmp[(size_t)k].v;
                dout[(int64_t)k * S + s] = tmp[(size_t)k].d;
            }
        }
    }
    return 0;
}

// Forward mode through the float64 line path (the line traversal and the
// line-integrated table, whose cells are linear in q): as oracle_line_samples,
// with dout the derivatives along tan.
int oracle_line_jvp(void* hv, int nS, const int* sslots, const unsigned char* sflags, int strategy, int64_t S,
                    const double* ox, const double* oy, const double* oz, const double* mx, const double* my,
                    const double* mz, const double* ha, const double* tan, double* out, double* dout, int nthreads) {
    Handle* H = static_cast<Handle*>(hv);
    if (H->precision != 64) {
        g_err = "oracle_line_jvp needs a float64 handle";
        return 1;
    }
    Catalog cat = make_catalog(nS, sslots, sflags, 0, nullptr, 0, nullptr, 0, nullptr);
    const Input<Dual> in = dual_input(H->d, tan);
    const LineLUT<Dual> lut(ph_lut_inorm);
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    {
        std::vector<Dual> tmp((size_t)(nS > 0 ? nS : 1));
#pragma omp for schedule(dynamic, 16)
        for (int64_t s = 0; s < S; ++s) {
            Dual o[3] = {Dual(ox[s]), Dual(oy[s]), Dual(oz[s])};
            Dual m[3] = {Dual(mx[s]), Dual(my[s]), Dual(mz[s])};
            Dual h = ha ? Dual(ha[s]) : Dual(0.0);
            line_interpolate<Dual>(in, lut, cat, strategy, o, m, h, tmp.data(), 1);
            for (int k = 0; k < nS; ++k) {
                out[(int64_t)k * S + s] = tmp[(size_t)k].v;
                dout[(int64_t)k * S + s] = tmp[(size_t)k].d;
            }
        }
    }
    return 0;
}

// query_leaves per point: counts, closest, closest d2; with pool (sized by the
// caller from the counts) the accepted indices point after point.
int oracle_query(void* hv, int64_t S, const double* px, const double* py, const double* pz, const double* radius,
                 int64_t* count, int64_t* closest, double* cd2, int64_t* pool) {
    Handle* H = static_cast<Handle*>(hv);
    auto run = [&](auto& in) {
        using T = std::decay_t<decltype(in.x[0])>;
        int64_t off = 0;
        for (int64_t s = 0; s < S; ++s) {
            T p[3] = {T(px[s]), T(py[s]), T(pz[s])};
            std::vector<i64> v;
            i64 best = 0;
            T bd = T(0);
            count[s] = query_leaves<T>(in.bvh, p, T(radius[s]), pool ? &v : nullptr, &best, &bd);
            closest[s] = best;
            cd2[s] = double(bd);
            if (pool)
                for (i64 k : v) pool[off++] = k;
        }
    };
    if (H->precision == 64) run(H->d);
    else run(H->f);
    return 0;
}

// LBVH_find_nearest_h per point; best (1-based leaf index) optional.
int oracle_nearest_h(void* hv, int64_t S, const double* px, const double* py, const double* pz, double* ha,
                     int64_t* best) {
    Handle* H = static_cast<Handle*>(hv);
    auto run = [&](auto& in) {
        using T = std::decay_t<decltype(in.x[0])>;
        for (int64_t s = 0; s < S; ++s) {
            T p[3] = {T(px[s]), T(py[s]), T(pz[s])};
            int64_t b = 0;
            ha[s] = double(nearest_h<T>(in.bvh, p, &b));
            if (best) best[s] = b;
        }
    };
    if (H->precision == 64) run(H->d);
    else run(H->f);
    return 0;
}

// Accepted leaves of one point (mode 0) or line (mode 1) traversal, in DFS
// order; returns the count (entries beyond cap are counted, not written).
int64_t oracle_neighbors(void* hv, int mode, int strategy, const double* a, const double* dir, double ha,
                         int64_t* idx, double* d2out, int64_t cap) {
    Handle* H = static_cast<Handle*>(hv);
    int64_t cnt = 0;
    auto run = [&](auto& in) {
        using T = std::decay_t<decltype(in.x[0])>;
        T Kv = kernel_support<T>(in.kernel);
        T p[3] = {T(a[0]), T(a[1]), T(a[2])};
        auto hit = [&](i64 k, T d2, T) {
            if (cnt < cap) { idx[cnt] = k; if (d2out) d2out[cnt] = double(d2); }
            ++cnt;
        };
        if (mode == 0) point_traversal<T>(in.bvh, p, strategy, Kv, T(ha), hit);
        else {
            T m[3] = {T(dir[0]), T(dir[1]), T(dir[2])};
            line_traversal<T>(in.bvh, p, m, strategy, Kv, T(ha), hit);
        }
    };
    if (H->precision == 64) run(H->d);
    else run(H->f);
    return cnt;
}

// ---- kernel-function and Morton building blocks (float64) ----
double oracle_kernel_f(int K, double q) { return kernel_f<double>(K, q); }
double oracle_kernel_df(int K, double q) { return kernel_df<double>(K, q); }
double oracle_kernel_sigma3(int K) { return kernel_sigma3<double>(K); }
double oracle_kernel_support(int K) { return kernel_support<double>(K); }
double oracle_kernel_W(int K, double r, double h) { return kernel_W<double>(K, r, h); }
void oracle_kernel_gradW(int K, double dx, double dy, double dz, double h, double* g) {
    kernel_gradW<double>(K, dx, dy, dz, h, g);
}
double oracle_line_W(int K, double r, double h) {
    static LineLUT<double> lut(ph_lut_inorm);
    return lut.W(K, r, h);
}
float oracle_line_W32(int K, float r, float h) {
    static LineLUT<float> lut(ph_lut_inorm);
    return lut.W(K, r, h);
}
uint32_t oracle_expand_bits3_u32(uint32_t x) { return expand_bits3(x); }
uint64_t oracle_expand_bits3_u64(uint64_t x) { return expand_bits3(x); }
uint32_t oracle_expand_bits2_u32(uint32_t x) { return expand_bits2(x); }
uint64_t oracle_expand_bits2_u64(uint64_t x) { return expand_bits2(x); }

// _find_range / _split_position over a caller-provided sorted code array.
void oracle_find_range_u64(const uint64_t* codes, int64_t n, int64_t i, int64_t* lo, int64_t* hi) {
    std::vector<u64> c(codes, codes + n);
    find_range(c, i, *lo, *hi);
}
int64_t oracle_split_position_u64(const uint64_t* codes, int64_t n, int64_t first, int64_t last) {
    std::vector<u64> c(codes, codes + n);
    return split_position(c, first, last);
}

}  // extern "C"
