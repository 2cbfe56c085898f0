// Implementation of the C++ drop-in API (include/bem2d/) over the C-ABI (include/bem2d_hip.h).
// Host code only: flattens geometries into the C-ABI's element format, moves matrices between the
// reference's host Matrix and device buffers, and maps C-ABI statuses onto the reference's
// exception tags.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <cstring>
#include <mutex>

#include "../../include/bem2d/bem2d.h"
#include "../../include/bem2d_hip.h"

namespace bem2d {

// ---- status mapping --------------------------------------------------------------------------
static void raise(int rc)
{
    if (rc == BEM2D_OK) return;
    if (std::getenv("BEM2D_VERBOSE")) std::fprintf(stderr, "bem2d: %s\n", bem2d_last_error());
    switch (rc) {
    case BEM2D_ERR_ARG:
    case BEM2D_ERR_UNSUPPORTED: throw ParameterException();
    case BEM2D_ERR_SINGULAR: throw LapackError();
    default: throw DeviceError();
    }
}

// ---- special functions (lib/bem2d_mathroutines.cpp:20-75) -------------------------------------
complex BesselH0(double x)
{
    double v[4];
    raise(bem2d_hankel01_host(&x, 1, v));
    return complex(v[0], v[1]);
}
complex BesselH1(double x)
{
    double v[4];
    raise(bem2d_hankel01_host(&x, 1, v));
    return complex(v[2], v[3]);
}
// BesselH0/H1(complex) (lib/bem2d_mathroutines.cpp:44-75, AMOS zbesh there) on the library's
// complex Hankel routine (integration path chosen by arg z).
static void hankel_complex(complex z, double* v)
{
    if (!(z.real() > 0)) throw ParameterException();
    const double zz[2] = {z.real(), z.imag()};
    raise(bem2d_hankel01_complex_host(std::arg(z), zz, 1, v));
}
complex BesselH0(complex z)
{
    double v[4];
    hankel_complex(z, v);
    return complex(v[0], v[1]);
}
complex BesselH1(complex z)
{
    double v[4];
    hankel_complex(z, v);
    return complex(v[2], v[3]);
}

// ---- Curve::Length (lib/bem2d_curve.cpp:11-25) ---------------------------------------------
// Curves the library knows: its adaptive Gauss-Legendre rule (rounding-level accurate); other
// curves: composite 16-point Gauss-Legendre on 64 panels of the virtual Deriv.
double Curve::Length(double t1, double t2) const
{
    int type = -1;
    double args[4] = {0, 0, 0, 0};
    if (device_desc(type, args) && t1 <= t2) {
        double out = 0;
        raise(bem2d_curve_length(type, args, t1, t2, &out));
        return out;
    }
    const int panels = 64, q = 16;
    double x[q], w[q];
    bem2d_gauss_legendre(q, x, w);
    double h = (t2 - t1) / panels, s = 0;
    for (int p = 0; p < panels; p++)
        for (int i = 0; i < q; i++) s += w[i] * length(Deriv(t1 + (p + x[i]) * h));
    return s * h;
}

void Gauss(dvector& x, dvector& w, int N)
{
    if (N < 1) throw ParameterException();
    x.resize(N);
    w.resize(N);
    raise(bem2d_gauss_legendre(N, x.data(), w.data()));
}

// ---- dense algebra (lib/bem2d_mathroutines.cpp:143-206) ---------------------------------------
Matrix operator+(const Matrix& lhs, const Matrix& rhs)
{
    if (lhs.dim[0] != rhs.dim[0] || lhs.dim[1] != rhs.dim[1]) throw ArrayMismatch();
    Matrix r(rhs);
    for (std::size_t i = 0; i < r.data->size(); i++) (*r.data)[i] += (*lhs.data)[i];
    return r;
}
Matrix operator-(const Matrix& lhs, const Matrix& rhs)
{
    if (lhs.dim[0] != rhs.dim[0] || lhs.dim[1] != rhs.dim[1]) throw ArrayMismatch();
    Matrix r(lhs);
    for (std::size_t i = 0; i < r.data->size(); i++) (*r.data)[i] -= (*rhs.data)[i];
    return r;
}
Matrix operator*(const Matrix& lhs, const complex& alpha)
{
    Matrix r(lhs);
    for (auto& v : *r.data) v *= alpha;
    return r;
}
Matrix operator*(const complex& alpha, const Matrix& rhs) { return rhs * alpha; }
Matrix operator*(const Matrix& lhs, const double& alpha)
{
    Matrix r(lhs);
    for (auto& v : *r.data) v *= alpha;
    return r;
}
Matrix operator*(const double& alpha, const Matrix& rhs) { return rhs * alpha; }
Matrix operator*(const Matrix& lhs, const Matrix& rhs)
{
    if (lhs.dim[1] != rhs.dim[0]) throw ArrayMismatch();
    const int m = lhs.dim[0], n = rhs.dim[1], kk = lhs.dim[1];
    Matrix r(m, n);
    for (int j = 0; j < n; j++)
        for (int p = 0; p < kk; p++) {
            const complex b = (*rhs.data)[p + (std::size_t)kk * j];
            for (int i = 0; i < m; i++)
                (*r.data)[i + (std::size_t)m * j] += (*lhs.data)[i + (std::size_t)m * p] * b;
        }
    return r;
}

namespace device {

// ---- process-wide device context ---------------------------------------------------------------
static bem2d_ctx* context()
{
    static std::once_flag once;
    static bem2d_ctx* ctx = nullptr;
    static int rc = BEM2D_OK;
    std::call_once(once, [] {
        const char* d = std::getenv("BEM2D_DEVICE");
        rc = bem2d_ctx_create(&ctx, d ? std::atoi(d) : 0);
    });
    raise(rc);
    return ctx;
}

// Device buffer owned for the duration of one call.
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(std::size_t bytes) { raise(bem2d_device_alloc(context(), (int64_t)bytes, &p)); }
    ~DevBuf() { bem2d_device_free(context(), p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// Flattened geometry + its bem2d_mesh: n elements, each carrying the same nb bases (the FlatMap
// rows / columns are element * nb + basis).
struct MeshHandle {
    bem2d_mesh* mesh = nullptr;
    int64_t n = 0;
    std::vector<bem2d_basis> bases;
    int nb() const { return (int)bases.size(); }
    int64_t size() const { return n * nb(); }
    // one ConstBasis or PolBasis(0) per element: the P0 kernels
    bool p0() const
    {
        return bases.size() == 1 && (bases[0].kind == BEM2D_BASIS_CONST ||
                                     (bases[0].kind == BEM2D_BASIS_POL && bases[0].degree == 0));
    }
    ~MeshHandle()
    {
        if (mesh) bem2d_mesh_destroy(mesh);
    }
};

static bool same_basis(const bem2d_basis& a, const bem2d_basis& b)
{
    return a.kind == b.kind && a.degree == b.degree && a.direction == b.direction &&
           a.k_re == b.k_re && a.k_im == b.k_im;
}

// The FlatMap must be element-major with the same basis list on every element (what AddBasis on
// the whole geometry produces); other layouts throw ParameterException.
static void flatten(const Geometry::flat_basis_map& flat, QuadOption opts, MeshHandle& h)
{
    if (flat.empty()) throw ParameterException();
    std::vector<int32_t> type;
    std::vector<double> param;
    std::vector<int64_t> index, next, prev;
    for (std::size_t i = 0; i < flat.size();) {
        const pElement& e = flat[i].first;
        std::vector<bem2d_basis> list;
        std::size_t j = i;
        for (; j < flat.size() && flat[j].first == e; ++j) {
            bem2d_basis d;
            if (!flat[j].second || !flat[j].second->device_desc(d)) throw ParameterException();
            list.push_back(d);
        }
        if (type.empty()) {
            if (list.size() > BEM2D_MAX_BASES) throw ParameterException();
            h.bases = list;
        } else if (list.size() != h.bases.size() ||
                   !std::equal(list.begin(), list.end(), h.bases.begin(), same_basis)) {
            throw ParameterException();
        }
        int t = -1;
        param.resize(param.size() + 8);
        if (!e->device_desc(t, &param[param.size() - 8])) throw ParameterException();
        type.push_back(t);
        index.push_back((int64_t)e->index());
        next.push_back(e->next() == (std::size_t)-1 ? -1 : (int64_t)e->next());
        prev.push_back(e->prev() == (std::size_t)-1 ? -1 : (int64_t)e->prev());
        i = j;
    }
    const int64_t n = (int64_t)type.size();
    bem2d_geometry g{n, type.data(), param.data(), index.data(), next.data(), prev.data()};
    bem2d_quad_option q{opts.L, opts.N, opts.sigma};
    raise(bem2d_mesh_create(context(), &g, q, &h.mesh));
    h.n = n;
}

static void flatten(const Geometry& geom, QuadOption opts, MeshHandle& h)
{
    flatten(*geom.FlatMap(), opts, h);
}

static pMatrix assemble(const Geometry& geom, QuadOption opts, int kind, freqtype k, double eta,
                        double ms, double ks)
{
    MeshHandle h;
    flatten(geom, opts, h);
    const int64_t n = h.size();
    pMatrix out(new Matrix((int)n));
    DevBuf A(sizeof(complex) * n * n);
    // k.im != 0: the kernels' complex-k branch (lib/bem2d_kernel.cpp:22-26 and alike)
    if (h.p0())
        raise(bem2d_assemble_complex(context(), h.mesh, kind, k.re, k.im, eta, ms, ks, A.p, n));
    else
        raise(bem2d_assemble_basis(context(), h.mesh, h.bases.data(), h.nb(), kind, k.re, k.im,
                                   eta, ms, ks, A.p, n));
    raise(bem2d_copy_to_host(context(), out->data->data(), A.p, sizeof(complex) * n * n));
    return out;
}

pMatrix AssembleCombined(const Geometry& geom, QuadOption opts, freqtype k, double eta,
                         double mass_scale, double kernel_scale)
{
    return assemble(geom, opts, BEM2D_KERNEL_COMBINED, k, eta, mass_scale, kernel_scale);
}

static pMatrix rhs(const Geometry& geom, QuadOption opts, int kind, freqtype k, double eta,
                   Point dir)
{
    MeshHandle h;
    flatten(geom, opts, h);
    const int64_t n = h.size();
    pMatrix out(new Matrix((int)n, 1));
    DevBuf B(sizeof(complex) * n);
    double d[2] = {dir.x, dir.y};
    if (h.p0())
        raise(bem2d_rhs_complex(context(), h.mesh, kind, k.re, k.im, eta, d, 1, 1.0, B.p, n));
    else
        raise(bem2d_rhs_basis(context(), h.mesh, h.bases.data(), h.nb(), kind, k.re, k.im, eta, d,
                              1, 1.0, B.p, n));
    raise(bem2d_copy_to_host(context(), out->data->data(), B.p, sizeof(complex) * n));
    return out;
}

} // namespace device

pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, SingleLayer g)
{
    return device::assemble(geom, opts, BEM2D_KERNEL_SL, g.k(), 0.0, 0.0, 1.0);
}
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, DoubleLayer g)
{
    return device::assemble(geom, opts, BEM2D_KERNEL_DL, g.k(), 0.0, 0.0, 1.0);
}
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, ConjDoubleLayer g)
{
    return device::assemble(geom, opts, BEM2D_KERNEL_CDL, g.k(), 0.0, 0.0, 1.0);
}
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, CombinedSingleConjDouble g)
{
    return device::assemble(geom, opts, BEM2D_KERNEL_COMBINED, g.getk(), g.eta(), 0.0, 1.0);
}
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, IdKernel)
{
    return device::assemble(geom, opts, BEM2D_KERNEL_ID, freqtype{1.0, 0.0}, 0.0, 0.0, 1.0);
}
pMatrix EvalIdent(const Geometry& geom, QuadOption opts)
{
    return device::assemble(geom, opts, BEM2D_KERNEL_NONE, freqtype{1.0, 0.0}, 0.0, 1.0, 0.0);
}

pMatrix DotProdBasFuns(const Geometry& geom, QuadOption ops, PlaneWave fun)
{
    return device::rhs(geom, ops, BEM2D_FIELD_PLANE, fun.k(), 0.0, fun.direction());
}
pMatrix DotProdBasFuns(const Geometry& geom, QuadOption ops, NormalPlaneWave fun)
{
    return device::rhs(geom, ops, BEM2D_FIELD_NORMAL_PLANE, fun.k(), 0.0, fun.direction());
}
pMatrix DotProdBasFuns(const Geometry& geom, QuadOption ops, CombinedPlaneWave fun)
{
    return device::rhs(geom, ops, BEM2D_FIELD_COMBINED_PLANE, fun.k(), fun.eta(), fun.direction());
}

void EvalKernel(const Geometry::flat_basis_map& basfuns, const std::vector<Point>& points,
                cvector& vals, Matrix& alpha, SingleLayer g, const Gauss1D& g1d)
{
    if (points.size() != vals.size()) throw SizeError();
    QuadOption opts{1, g1d.Size(), 0.5};
    device::MeshHandle h;
    device::flatten(basfuns, opts, h);
    const int64_t n = h.size(), np = (int64_t)points.size();
    if ((int64_t)alpha.data->size() < n) throw SizeError();
    if (np == 0) return;
    device::DevBuf c(sizeof(complex) * n), P(sizeof(double) * 2 * np), out(sizeof(complex) * np);
    std::vector<double> pts(2 * np);
    for (int64_t i = 0; i < np; i++) {
        pts[2 * i] = points[i].x;
        pts[2 * i + 1] = points[i].y;
    }
    raise(bem2d_copy_to_device(device::context(), c.p, alpha.data->data(), sizeof(complex) * n));
    raise(bem2d_copy_to_device(device::context(), P.p, pts.data(), sizeof(double) * 2 * np));
    if (h.p0())
        raise(bem2d_eval_single_layer_complex(device::context(), h.mesh, g.k().re, g.k().im, c.p,
                                              (const double*)P.p, np, out.p));
    else
        raise(bem2d_eval_single_layer_basis(device::context(), h.mesh, h.bases.data(), h.nb(),
                                            g.k().re, g.k().im, c.p, (const double*)P.p, np,
                                            out.p));
    cvector v(np);
    raise(bem2d_copy_to_host(device::context(), v.data(), out.p, sizeof(complex) * np));
    for (int64_t i = 0; i < np; i++) vals[i] += v[i];
}

pMatrix FarField(const Geometry& geom, QuadOption opts, freqtype k, Matrix& alpha,
                 const dvector& theta)
{
    if (k.im != 0) throw ParameterException();
    device::MeshHandle h;
    device::flatten(geom, opts, h);
    const int64_t n = h.size(), na = (int64_t)theta.size();
    const int nrhs = alpha.dim[1];
    if (alpha.dim[0] != n || na == 0) throw SizeError();
    pMatrix out(new Matrix((int)na, nrhs));
    device::DevBuf c(sizeof(complex) * n * nrhs), th(sizeof(double) * na),
        o(sizeof(complex) * na * nrhs);
    raise(bem2d_copy_to_device(device::context(), c.p, alpha.data->data(), sizeof(complex) * n * nrhs));
    raise(bem2d_copy_to_device(device::context(), th.p, theta.data(), sizeof(double) * na));
    if (h.p0())
        raise(bem2d_farfield(device::context(), h.mesh, k.re, c.p, n, nrhs, (const double*)th.p, na,
                             o.p));
    else
        raise(bem2d_farfield_basis(device::context(), h.mesh, h.bases.data(), h.nb(), k.re, c.p, n,
                                   nrhs, (const double*)th.p, na, o.p));
    raise(bem2d_copy_to_host(device::context(), out->data->data(), o.p, sizeof(complex) * na * nrhs));
    return out;
}

// ---- point in polygon (InPolygon, lib/bem2d_mathroutines.cpp:732-758) -------------------------
// Crossing-number test: a horizontal ray from the point crosses the closed polygon through the
// vertices an odd number of times iff the point is inside.
static int inside(const PointVector& v, const Point& p)
{
    int c = 0;
    const std::size_t m = v.size();
    for (std::size_t i = 0, j = m - 1; i < m; j = i++) {
        const bool straddle = (v[i].y > p.y) != (v[j].y > p.y);
        if (straddle && p.x < (v[j].x - v[i].x) * (p.y - v[i].y) / (v[j].y - v[i].y) + v[i].x)
            c = !c;
    }
    return c;
}

void InPolygon(const std::vector<pGeometry> polygons, const std::vector<Point>& testpoints,
               std::vector<int>& inpoly)
{
    inpoly.assign(testpoints.size(), 0);
    for (const pGeometry& g : polygons) {
        const PointVector& v = g->points();
        if (v.empty()) continue;
        for (std::size_t j = 0; j < testpoints.size(); ++j) inpoly[j] += inside(v, testpoints[j]);
    }
}

void InPolygon(const pGeometry polygon, const std::vector<Point>& testpoints,
               std::vector<int>& inpoly)
{
    InPolygon(std::vector<pGeometry>{polygon}, testpoints, inpoly);
}

// ---- AnalyticCurve (lib/bem2d_shape.cpp:247-331) -----------------------------------------------
static void curve_desc(const pCurve& curve, int& type, double* args)
{
    if (!curve || !curve->device_desc(type, args)) throw ParameterException();
}

AnalyticCurve::AnalyticCurve(int n, pCurve curve, int closed) : curve_(curve)
{
    if (n < 2) throw ParameterException();
    const double curveLength = curve_->Length();
    dvector partition(n + 1);
    for (int i = 0; i <= n; i++) partition[i] = i * curveLength / n;
    ParameterizeArc(partition);
    build(closed);
}

AnalyticCurve::AnalyticCurve(int ppw, freqtype k, pCurve curve, int closed, int L, double sigma)
    : curve_(curve)
{
    const double curveLength = curve_->Length();
    dvector partition;
    partition.push_back(0);
    int c = (int)std::ceil(ppw * (double)k.re * curveLength / 2.0 / PI);
    c = std::max(c, 10);
    // graded towards both ends: L pieces of ratio sigma inside the first and the last interval
    if (L == 0) {
        partition.push_back(curveLength / c);
    } else {
        double tsigma = 1;
        for (int m = 0; m < L; m++) tsigma = tsigma * sigma;
        partition.push_back(tsigma * curveLength / c);
        for (int m = 0; m < L; m++) {
            tsigma = tsigma / sigma;
            partition.push_back(tsigma * curveLength / c);
        }
    }
    for (int j = 2; j < c; j++) partition.push_back(curveLength / c * j);
    if (L == 0) {
        partition.push_back(curveLength);
    } else {
        const double td = curveLength / c * (c - 1);
        double tsigma = 1;
        for (int m = 0; m < L; m++) {
            tsigma = tsigma * sigma;
            partition.push_back(td + curveLength / c * (1 - tsigma));
        }
        partition.push_back(curveLength);
    }
    ParameterizeArc(partition);
    build(closed);
}

void AnalyticCurve::ParameterizeArc(dvector& partition)
{
    const int n = (int)partition.size() - 1;
    if (n < 1) throw ParameterException();
    int type = -1;
    double args[4] = {0, 0, 0, 0};
    curve_desc(curve_, type, args);
    arclengthparam.assign(n + 1, 0.0);
    raise(bem2d_arclength_params(type, args, &partition[1], n, &arclengthparam[1]));
    arclengthparam[0] = 0;
    arclengthparam[n] = 1; // the curve's end, as the reference sets it
}

void AnalyticCurve::build(int closed)
{
    const int n = (int)arclengthparam.size() - 1;
    elements_.resize(n);
    for (int i = 0; i < n - 1; i++)
        elements_[i] = pElement(new AnalyticCurveElement(arclengthparam[i], arclengthparam[i + 1],
                                                         curve_, i));
    elements_[n - 1] = pElement(new AnalyticCurveElement(arclengthparam[n - 1], 1, curve_, n - 1));
    for (int i = 1; i < n - 1; i++) {
        elements_[i]->set_next(i + 1);
        elements_[i]->set_prev(i - 1);
    }
    elements_[0]->set_next(1);
    elements_[n - 1]->set_prev(n - 2);
    if (closed) {
        elements_[0]->set_prev(n - 1);
        elements_[n - 1]->set_next(0);
    }
}

pGeometry AnalyticCurve::GetGeometry() { return pGeometry(new Geometry(elements_)); }

// ---- spectral drivers (lib/bem2d_mathroutines.cpp:208-320, 353-410, 534-653) on the GPU ---------
static void square(const Matrix& m)
{
    if (m.dim[0] != m.dim[1]) throw ArrayMismatch();
}

void L2NormCond(const Matrix& stiff, double& norm, double& cond)
{
    square(stiff);
    const int64_t n = stiff.dim[0];
    device::DevBuf A(sizeof(complex) * n * n);
    raise(bem2d_copy_to_device(device::context(), A.p, stiff.data->data(), sizeof(complex) * n * n));
    raise(bem2d_l2_norm_cond(device::context(), n, A.p, n, &norm, &cond));
}

void SingularValues(const Matrix& stiff, pdvector& singvals)
{
    square(stiff);
    const int64_t n = stiff.dim[0];
    singvals = pdvector(new dvector(n));
    device::DevBuf A(sizeof(complex) * n * n), S(sizeof(double) * n);
    raise(bem2d_copy_to_device(device::context(), A.p, stiff.data->data(), sizeof(complex) * n * n));
    raise(bem2d_singular_values(device::context(), n, A.p, n, (double*)S.p));
    raise(bem2d_copy_to_host(device::context(), singvals->data(), S.p, sizeof(double) * n));
}

void HermitianEigenvalues(const Matrix& k, pdvector& evalues)
{
    square(k);
    const int64_t n = k.dim[0];
    evalues = pdvector(new dvector(n));
    device::DevBuf A(sizeof(complex) * n * n), W(sizeof(double) * n);
    raise(bem2d_copy_to_device(device::context(), A.p, k.data->data(), sizeof(complex) * n * n));
    raise(bem2d_hermitian_eigenvalues(device::context(), n, A.p, n, (double*)W.p));
    raise(bem2d_copy_to_host(device::context(), evalues->data(), W.p, sizeof(double) * n));
}

void Eigenvalues(const Matrix& k, pcvector& evalues)
{
    square(k);
    const int64_t n = k.dim[0];
    evalues = pcvector(new cvector(n));
    device::DevBuf A(sizeof(complex) * n * n), W(sizeof(complex) * n);
    raise(bem2d_copy_to_device(device::context(), A.p, k.data->data(), sizeof(complex) * n * n));
    const int rc = bem2d_eigenvalues(device::context(), n, A.p, n, W.p);
    if (rc == BEM2D_ERR_NOT_CONVERGED) throw LapackError(); // zgeev's info > 0
    raise(rc);
    raise(bem2d_copy_to_host(device::context(), evalues->data(), W.p, sizeof(complex) * n));
}

Matrix ChangeBasis(const Matrix& k, const Matrix& m)
{
    square(k);
    square(m);
    if (k.dim[0] != m.dim[0]) throw ArrayMismatch();
    const int64_t n = k.dim[0];
    Matrix out(k);
    device::DevBuf K(sizeof(complex) * n * n), M(sizeof(complex) * n * n);
    raise(bem2d_copy_to_device(device::context(), K.p, k.data->data(), sizeof(complex) * n * n));
    raise(bem2d_copy_to_device(device::context(), M.p, m.data->data(), sizeof(complex) * n * n));
    raise(bem2d_change_basis(device::context(), n, K.p, n, M.p, n));
    raise(bem2d_copy_to_host(device::context(), out.data->data(), K.p, sizeof(complex) * n * n));
    return out;
}

void NumRange(const Matrix& m, int n, std::string filename)
{
    square(m);
    const int64_t N = m.dim[0];
    std::vector<double> pts(2 * (2 * (std::size_t)n + 1));
    device::DevBuf A(sizeof(complex) * N * N);
    raise(bem2d_copy_to_device(device::context(), A.p, m.data->data(), sizeof(complex) * N * N));
    raise(bem2d_numerical_range(device::context(), N, A.p, N, n, pts.data()));
    std::ofstream out(filename.c_str());
    for (int i = 0; i < 2 * n + 1; ++i) out << pts[2 * i] << " " << pts[2 * i + 1] << std::endl;
}

Matrix ConjTranspose(const Matrix& m)
{
    const int M = m.dim[0], N = m.dim[1];
    Matrix r(N, M);
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i)
            (*r.data)[j + (std::size_t)N * i] = std::conj((*m.data)[i + (std::size_t)M * j]);
    return r;
}

pMatrix SolveSystem(Matrix& m, Matrix& rhs)
{
    const int64_t n = m.dim[0], nrhs = rhs.dim[1];
    if (m.dim[1] != n || rhs.dim[0] != n) throw ArrayMismatch();
    pMatrix res(new Matrix(rhs));
    device::DevBuf A(sizeof(complex) * n * n), B(sizeof(complex) * n * nrhs);
    raise(bem2d_copy_to_device(device::context(), A.p, m.data->data(), sizeof(complex) * n * n));
    raise(bem2d_copy_to_device(device::context(), B.p, res->data->data(), sizeof(complex) * n * nrhs));
    raise(bem2d_solve_fp64(device::context(), n, nrhs, A.p, n, B.p, n, 0));
    raise(bem2d_copy_to_host(device::context(), res->data->data(), B.p, sizeof(complex) * n * nrhs));
    return res;
}

} // namespace bem2d
