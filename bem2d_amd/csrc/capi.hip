// C-ABI of the MI355X hot path (declared in include/bem2d_hip.h): contexts, meshes, FP64 assembly,
// right-hand sides, the plain FP64 solve, the single-layer potential and the far field.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bem2d_hip.h"
#include "assembly_kernels.h"
#include "basis.h"
#include "internal.h"

using namespace bem2d_dev;

// ------------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------------
static thread_local std::string g_last_error;

int bem2d_internal_fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

extern "C" const char* bem2d_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------------------------------
extern "C" int bem2d_ctx_create(bem2d_ctx** out, int device)
{
    if (!out) return FAIL(BEM2D_ERR_ARG, "bem2d_ctx_create: null output");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return FAIL(BEM2D_ERR_ARG, "bem2d_ctx_create: device %d of %d", device, ndev);
    HIPCHK(hipSetDevice(device));
    bem2d_ctx* c = new bem2d_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return FAIL(BEM2D_ERR_DEVICE, "hipStreamCreate failed");
    }
    c->stream = c->own_stream;
    if (hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        rocblas_create_handle(&c->blas) != rocblas_status_success) {
        if (c->aux) (void)hipStreamDestroy(c->aux);
        if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
        if (c->ev_join) (void)hipEventDestroy(c->ev_join);
        (void)hipStreamDestroy(c->own_stream);
        delete c;
        return FAIL(BEM2D_ERR_DEVICE, "stream / event / rocBLAS handle creation failed");
    }
    rocblas_set_stream(c->blas, c->stream);
    // device address of the no-allocation solve's workspace, taken once per device while it is
    // current (a failure here is reported by that entry point, which tries again)
    (void)bem2d_internal_static_scratch(device, nullptr);
    // the autotuner's table: $BEM2D_TUNE_FILE, else tuned/<arch>.tune beside this library
    const char* tf = std::getenv("BEM2D_TUNE_FILE");
    if (tf && tf[0]) {
        if (bem2d_tune_load(c, tf) != BEM2D_OK) {
            bem2d_ctx_destroy(c);
            return FAIL(BEM2D_ERR_ARG, "BEM2D_TUNE_FILE %s: cannot use it", tf);
        }
    } else {
        Dl_info info;
        if (dladdr((void*)&bem2d_ctx_create, &info) && info.dli_fname) {
            std::string path(info.dli_fname);
            path = path.substr(0, path.find_last_of('/') + 1) + "tuned/gfx950.tune";
            if (FILE* f = std::fopen(path.c_str(), "r")) {
                std::fclose(f);
                if (bem2d_tune_load(c, path.c_str()) != BEM2D_OK) {
                    bem2d_ctx_destroy(c);
                    return FAIL(BEM2D_ERR_ARG, "tune table %s: cannot use it", path.c_str());
                }
            }
        }
    }
    *out = c;
    return BEM2D_OK;
}

// ------------------------------------------------------------------------------------------------
// launch-shape table (written by the autotuner, bem2d_amd/autotune.py)
// ------------------------------------------------------------------------------------------------
extern "C" int bem2d_tune_set_assembly(bem2d_ctx* c, int nbuckets, const int64_t* upper_tiles,
                                       const int* split)
{
    if (!c || nbuckets < 0 || nbuckets > kMaxTuneBuckets || (nbuckets > 0 && (!upper_tiles || !split)))
        return FAIL(BEM2D_ERR_ARG, "bem2d_tune_set_assembly: bad argument");
    for (int b = 0; b < nbuckets; ++b) {
        if (split[b] != 1 && split[b] != 2 && split[b] != 4)
            return FAIL(BEM2D_ERR_ARG, "bem2d_tune_set_assembly: split %d not in {1, 2, 4}", split[b]);
        if (upper_tiles[b] <= 0 || (b > 0 && upper_tiles[b] <= upper_tiles[b - 1]))
            return FAIL(BEM2D_ERR_ARG, "bem2d_tune_set_assembly: bucket bounds must increase");
    }
    c->tune.nasm = nbuckets;
    for (int b = 0; b < nbuckets; ++b) {
        c->tune.asm_upper[b] = upper_tiles[b];
        c->tune.asm_split[b] = split[b];
    }
    return BEM2D_OK;
}

extern "C" int bem2d_tune_get_assembly(const bem2d_ctx* c, int64_t* upper_tiles, int* split)
{
    if (!c) return -1;
    for (int b = 0; b < c->tune.nasm; ++b) {
        if (upper_tiles) upper_tiles[b] = c->tune.asm_upper[b];
        if (split) split[b] = c->tune.asm_split[b];
    }
    return c->tune.nasm;
}

extern "C" int bem2d_tune_set_lu_pair_tiles(bem2d_ctx* c, int tiles)
{
    if (!c || tiles < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_set_lu_pair_tiles: bad argument");
    c->tune.lu_pair_tiles = tiles;
    return BEM2D_OK;
}

extern "C" int bem2d_tune_get_lu_pair_tiles(const bem2d_ctx* c) { return c ? c->tune.lu_pair_tiles : -1; }

extern "C" int bem2d_tune_set_lu_quad_tiles(bem2d_ctx* c, int tiles)
{
    if (!c || tiles < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_set_lu_quad_tiles: bad argument");
    c->tune.lu_quad_tiles = tiles;
    return BEM2D_OK;
}

extern "C" int bem2d_tune_get_lu_quad_tiles(const bem2d_ctx* c) { return c ? c->tune.lu_quad_tiles : -1; }

extern "C" int bem2d_tune_set_lu_oct_tiles(bem2d_ctx* c, int tiles)
{
    if (!c || tiles < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_set_lu_oct_tiles: bad argument");
    c->tune.lu_oct_tiles = tiles;
    return BEM2D_OK;
}

extern "C" int bem2d_tune_get_lu_oct_tiles(const bem2d_ctx* c) { return c ? c->tune.lu_oct_tiles : -1; }

extern "C" int bem2d_tune_set_lu_split_tiles(bem2d_ctx* c, int tiles)
{
    if (!c || tiles < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_set_lu_split_tiles: bad argument");
    c->tune.lu_split_tiles = tiles;
    return BEM2D_OK;
}

extern "C" int bem2d_tune_get_lu_split_tiles(const bem2d_ctx* c) { return c ? c->tune.lu_split_tiles : -1; }

// Text table, one setting per line ('#' starts a comment):
//   asm_split <upper_tiles> <wavefronts per tile>     (buckets in increasing order)
//   lu_pair_tiles <tiles>
//   lu_quad_tiles <tiles>
//   lu_oct_tiles <tiles>
//   lu_split_tiles <tiles>
extern "C" int bem2d_tune_load(bem2d_ctx* c, const char* path)
{
    if (!c || !path) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_load: null argument");
    FILE* f = std::fopen(path, "r");
    if (!f) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_load: cannot open %s", path);
    int64_t upper[kMaxTuneBuckets];
    int split[kMaxTuneBuckets];
    int nb = 0, pair = c->tune.lu_pair_tiles, quad = c->tune.lu_quad_tiles;
    int oct = c->tune.lu_oct_tiles;
    int lusplit = c->tune.lu_split_tiles, rc = BEM2D_OK;
    char line[256];
    while (rc == BEM2D_OK && std::fgets(line, sizeof line, f)) {
        if (char* h = std::strchr(line, '#')) *h = 0;
        char key[64];
        long long a = 0, b = 0;
        const int got = std::sscanf(line, "%63s %lld %lld", key, &a, &b);
        if (got <= 0) continue;
        if (!std::strcmp(key, "asm_split") && got == 3 && nb < kMaxTuneBuckets) {
            upper[nb] = a;
            split[nb++] = (int)b;
        } else if (!std::strcmp(key, "lu_pair_tiles") && got == 2) {
            pair = (int)a;
        } else if (!std::strcmp(key, "lu_quad_tiles") && got == 2) {
            quad = (int)a;
        } else if (!std::strcmp(key, "lu_oct_tiles") && got == 2) {
            oct = (int)a;
        } else if (!std::strcmp(key, "lu_split_tiles") && got == 2) {
            lusplit = (int)a;
        } else {
            rc = FAIL(BEM2D_ERR_ARG, "bem2d_tune_load: bad line in %s: %s", path, line);
        }
    }
    std::fclose(f);
    if (rc) return rc;
    if (pair < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_load: lu_pair_tiles %d < 1 in %s", pair, path);
    if (quad < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_load: lu_quad_tiles %d < 1 in %s", quad, path);
    if (oct < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_tune_load: lu_oct_tiles %d < 1 in %s", oct, path);
    if (lusplit < 1)
        return FAIL(BEM2D_ERR_ARG, "bem2d_tune_load: lu_split_tiles %d < 1 in %s", lusplit, path);
    if ((rc = bem2d_tune_set_assembly(c, nb, upper, split))) return rc;
    if ((rc = bem2d_tune_set_lu_split_tiles(c, lusplit))) return rc;
    if ((rc = bem2d_tune_set_lu_quad_tiles(c, quad))) return rc;
    if ((rc = bem2d_tune_set_lu_oct_tiles(c, oct))) return rc;
    return bem2d_tune_set_lu_pair_tiles(c, pair);
}

extern "C" void bem2d_ctx_destroy(bem2d_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->prof)
        for (hipEvent_t e : p.pending) (void)hipEventDestroy(e);
    for (auto& s : c->scratch)
        if (s.ptr) (void)hipFree(s.ptr);
    if (c->capture) (void)hipStreamDestroy(c->capture);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->ev_stage) (void)hipEventDestroy(c->ev_stage);
    bem2d_internal_mixed_release(c);
    bem2d_internal_rulesets_release(c);
    rocblas_destroy_handle(c->blas);
    (void)hipStreamSynchronize(c->aux);
    (void)hipStreamDestroy(c->aux);
    (void)hipEventDestroy(c->ev_fork);
    (void)hipEventDestroy(c->ev_join);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
}

extern "C" int bem2d_ctx_set_stream(bem2d_ctx* c, void* stream)
{
    if (!c) return FAIL(BEM2D_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->stream = (hipStream_t)stream;
    rocblas_set_stream(c->blas, c->stream);
    return BEM2D_OK;
}

extern "C" int bem2d_ctx_reset_stream(bem2d_ctx* c)
{
    if (!c) return FAIL(BEM2D_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->stream = c->own_stream;
    rocblas_set_stream(c->blas, c->stream);
    return BEM2D_OK;
}

extern "C" void* bem2d_ctx_stream(bem2d_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int bem2d_ctx_release_scratch(bem2d_ctx* c)
{
    if (!c) return FAIL(BEM2D_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& s : c->scratch) {
        if (s.ptr) HIPCHK(hipFree(s.ptr));
        s.ptr = nullptr;
        s.bytes = 0;
    }
    return BEM2D_OK;
}

extern "C" int bem2d_synchronize(bem2d_ctx* c)
{
    if (!c) return FAIL(BEM2D_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BEM2D_OK;
}

// ------------------------------------------------------------------------------------------------
// kernel timing
// ------------------------------------------------------------------------------------------------
void bem2d_internal_prof_begin(bem2d_ctx* c, int id, hipStream_t s)
{
    if (!c->profiling) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s ? s : c->stream);
    c->prof[id].pending.push_back(e);
}

void bem2d_internal_prof_end(bem2d_ctx* c, int id, hipStream_t s)
{
    if (!c->profiling || c->prof[id].pending.size() % 2 == 0) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s ? s : c->stream);
    c->prof[id].pending.push_back(e);
}

extern "C" int bem2d_profile_enable(bem2d_ctx* c, int on)
{
    if (!c) return FAIL(BEM2D_ERR_ARG, "null context");
    c->profiling = on != 0;
    return BEM2D_OK;
}

extern "C" int bem2d_profile_reset(bem2d_ctx* c)
{
    if (!c) return FAIL(BEM2D_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& p : c->prof) {
        for (hipEvent_t e : p.pending) (void)hipEventDestroy(e);
        p.pending.clear();
        p.total_ms = 0.0;
        p.launches = 0;
    }
    return BEM2D_OK;
}

extern "C" int bem2d_profile_read(bem2d_ctx* c, int id, double* total_ms, int64_t* launches)
{
    if (!c || id < 0 || id >= BEM2D_PROF_COUNT) return FAIL(BEM2D_ERR_ARG, "bem2d_profile_read: bad argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    auto& p = c->prof[id];
    const size_t pairs = p.pending.size() / 2;
    for (size_t q = 0; q < pairs; ++q) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, p.pending[2 * q], p.pending[2 * q + 1]));
        p.total_ms += ms;
        p.launches += 1;
        (void)hipEventDestroy(p.pending[2 * q]);
        (void)hipEventDestroy(p.pending[2 * q + 1]);
    }
    p.pending.erase(p.pending.begin(), p.pending.begin() + 2 * pairs);
    if (total_ms) *total_ms = p.total_ms;
    if (launches) *launches = p.launches;
    return BEM2D_OK;
}

void* bem2d_internal_scratch(bem2d_ctx* c, int slot, size_t bytes)
{
    auto& s = c->scratch[slot];
    if (s.bytes >= bytes && s.ptr) return s.ptr;
    if (s.ptr) {
        (void)hipStreamSynchronize(c->stream); // earlier work may still read the old buffer
        (void)hipFree(s.ptr);
        s.ptr = nullptr;
        s.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&s.ptr, want) != hipSuccess) {
        s.ptr = nullptr;
        return nullptr;
    }
    s.bytes = want;
    return s.ptr;
}

extern "C" int bem2d_device_alloc(bem2d_ctx* c, int64_t bytes, void** out)
{
    if (!c || !out || bytes < 0) return FAIL(BEM2D_ERR_ARG, "bem2d_device_alloc: bad argument");
    HIPCHK(hipSetDevice(c->device));
    if (hipMalloc(out, bytes > 0 ? (size_t)bytes : 1) != hipSuccess)
        return FAIL(BEM2D_ERR_NOMEM, "hipMalloc(%lld) failed", (long long)bytes);
    return BEM2D_OK;
}

extern "C" int bem2d_device_free(bem2d_ctx* c, void* p)
{
    if (!c) return FAIL(BEM2D_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (p) HIPCHK(hipFree(p));
    return BEM2D_OK;
}

extern "C" int bem2d_copy_to_device(bem2d_ctx* c, void* dst, const void* src, int64_t bytes)
{
    if (!c || bytes < 0) return FAIL(BEM2D_ERR_ARG, "bem2d_copy_to_device: bad argument");
    if (bytes == 0) return BEM2D_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BEM2D_OK;
}

extern "C" int bem2d_copy_to_host(bem2d_ctx* c, void* dst, const void* src, int64_t bytes)
{
    if (!c || bytes < 0) return FAIL(BEM2D_ERR_ARG, "bem2d_copy_to_host: bad argument");
    if (bytes == 0) return BEM2D_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BEM2D_OK;
}

// ------------------------------------------------------------------------------------------------
// quadrature rules on the host
// ------------------------------------------------------------------------------------------------
// Gauss-Legendre nodes (ascending) and weights on [0, 1]: Newton's method on P_n in long double.
// The reference obtains the same rule by Golub-Welsch (lib/bem2d_quadrature.cpp:10-56); both give
// the rule to within an ulp.
void bem2d_internal_gauss01(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    if (n == 1) {
        x[0] = 0.5;
        w[0] = 1.0;
        return;
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < n; ++i) {
        long double z = cosl(pi * (i + 0.75L) / (n + 0.5L)); // i-th largest root
        long double dp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int m = 2; m <= n; ++m) {
                long double p2 = ((2 * m - 1) * z * p1 - (m - 1) * p0) / m;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0L);
            long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        {
            long double p0 = 1.0L, p1 = z;
            for (int m = 2; m <= n; ++m) {
                long double p2 = ((2 * m - 1) * z * p1 - (m - 1) * p0) / m;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0L);
        }
        const int o = n - 1 - i; // ascending order
        x[o] = (double)((1.0L + z) / 2.0L);
        w[o] = (double)(1.0L / ((1.0L - z * z) * dp * dp)); // (2/((1-z^2)P'^2)) / 2
    }
}

namespace {

struct Rule {
    std::vector<double> x, y, w;
};

// Tensor rule of Gauss2D(N) rescaled onto [a,b] x [c,d] and appended (lib/bem2d_quadpoints.cpp:
// 166-204 and MapPoints2d, lib/bem2d_quadrature.cpp:68-76).
void append_box(Rule& r, const std::vector<double>& g, const std::vector<double>& gw, double a,
                double b, double c, double d)
{
    const int N = (int)g.size();
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            r.x.push_back(a + (b - a) * g[i]);
            r.y.push_back(c + (d - c) * g[j]);
            r.w.push_back((b - a) * (d - c) * (gw[i] * gw[j]));
        }
}

// AdaptedGauss1: L x L graded boxes on [sigma^L, 1]^2 with breakpoints sigma^L, ..., sigma, 1
// (lib/bem2d_quadpoints.cpp:210-235).  The strips below sigma^L are not covered, as in the reference.
Rule adapted1(const std::vector<double>& g, const std::vector<double>& gw, int L, double sigma)
{
    Rule r;
    std::vector<double> br;
    for (int i = 0; i < L; ++i) br.push_back(std::pow(sigma, (double)(L - i)));
    br.push_back(1.0);
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) append_box(r, g, gw, br[i], br[i + 1], br[j], br[j + 1]);
    return r;
}

// AdaptedGauss2: three boxes per level graded toward the corner (0, 1); AdaptedGauss3: toward (1, 0)
// (lib/bem2d_quadpoints.cpp:240-318).
Rule adapted23(const std::vector<double>& g, const std::vector<double>& gw, int L, double sigma,
               bool toward_10)
{
    Rule r;
    if (!toward_10) {
        double a = sigma, b = 1.0;
        for (int l = 0; l < L; ++l) {
            append_box(r, g, gw, 0, a, 1 - b, 1 - a);
            append_box(r, g, gw, a, b, 1 - b, 1 - a);
            append_box(r, g, gw, a, b, 1 - a, 1);
            a *= sigma;
            b *= sigma;
        }
    } else {
        double a = 1.0, b = sigma;
        for (int l = 0; l < L; ++l) {
            append_box(r, g, gw, 1 - a, 1 - b, 0, b);
            append_box(r, g, gw, 1 - a, 1 - b, b, a);
            append_box(r, g, gw, 1 - b, 1, b, a);
            a *= sigma;
            b *= sigma;
        }
    }
    return r;
}

// The abscissa block of a rule for the singular-pair kernel (layout at kRuleAbscissae): distinct
// x and y values (compared bit for bit, so a cached map is the one the node would compute) and
// every node's indices into them.
void append_abscissa_block(std::vector<double>& out, const Rule& r)
{
    const size_t m = r.w.size();
    std::vector<double> ux, uy, idx(2 * m);
    std::unordered_map<uint64_t, int> fx, fy;
    auto find = [](std::unordered_map<uint64_t, int>& f, std::vector<double>& u, double v) {
        uint64_t key;
        std::memcpy(&key, &v, sizeof key);
        auto it = f.find(key);
        if (it != f.end()) return it->second;
        f.emplace(key, (int)u.size());
        u.push_back(v);
        return (int)u.size() - 1;
    };
    for (size_t q = 0; q < m; ++q) {
        idx[2 * q] = find(fx, ux, r.x[q]);
        idx[2 * q + 1] = find(fy, uy, r.y[q]);
    }
    const bool fits = ux.size() <= (size_t)kRuleAbscissae && uy.size() <= (size_t)kRuleAbscissae;
    out.push_back(fits ? (double)ux.size() : 0.0);
    out.push_back(fits ? (double)uy.size() : 0.0);
    ux.resize(kRuleAbscissae, 0.0);
    uy.resize(kRuleAbscissae, 0.0);
    if (!fits) std::fill(idx.begin(), idx.end(), 0.0);
    out.insert(out.end(), ux.begin(), ux.end());
    out.insert(out.end(), uy.begin(), uy.end());
    out.insert(out.end(), idx.begin(), idx.end());
}

} // namespace

extern "C" int bem2d_gauss_legendre(int n, double* x, double* w)
{
    if (n < 1 || !x || !w) return FAIL(BEM2D_ERR_ARG, "bem2d_gauss_legendre: bad argument");
    std::vector<double> gx, gw;
    bem2d_internal_gauss01(n, gx, gw);
    std::memcpy(x, gx.data(), sizeof(double) * n);
    std::memcpy(w, gw.data(), sizeof(double) * n);
    return BEM2D_OK;
}

extern "C" int bem2d_quad_rule(int which, int N, int L, double sigma, double* x, double* y, double* w)
{
    if (N < 1 || L < 0 || !(sigma > 0 && sigma < 1) || which < 0 || which > 3) {
        FAIL(BEM2D_ERR_ARG, "bem2d_quad_rule: bad argument");
        return -1;
    }
    std::vector<double> gx, gw;
    bem2d_internal_gauss01(N, gx, gw);
    Rule r;
    if (which == 0) append_box(r, gx, gw, 0.0, 1.0, 0.0, 1.0);
    else if (which == 1) r = adapted1(gx, gw, L, sigma);
    else r = adapted23(gx, gw, L, sigma, which == 3);
    const int m = (int)r.w.size();
    if (x && y && w) {
        std::memcpy(x, r.x.data(), sizeof(double) * m);
        std::memcpy(y, r.y.data(), sizeof(double) * m);
        std::memcpy(w, r.w.data(), sizeof(double) * m);
    }
    return m;
}

// ------------------------------------------------------------------------------------------------
// mesh
// ------------------------------------------------------------------------------------------------
// Device bytes of a mesh of n elements (an upper bound: at most 3 singular entries per row).
size_t bem2d_internal_mesh_bytes(int64_t n, bem2d_quad_option q)
{
    const int64_t N = q.N, L = q.L;
    const int64_t nodes = N * N * L * L + 6 * N * N * L;
    return sizeof(int32_t) * n + sizeof(double) * 8 * n + sizeof(int64_t) * 3 * n +
           sizeof(double) * (NODE_PLANES * n * N + BOUND_PLANES * n) +
           sizeof(double) * (5 * nodes + 3 * (2 + 2 * (int64_t)kRuleAbscissae) + 1) +
           sizeof(double) * 2 * N + sizeof(int64_t) * (9 * n + 3) + 8 * 256;
}

extern "C" int bem2d_mesh_create(bem2d_ctx* c, const bem2d_geometry* g, bem2d_quad_option q,
                                 bem2d_mesh** out)
{
    if (!c) return FAIL(BEM2D_ERR_ARG, "bem2d_mesh_create: null argument");
    return bem2d_internal_mesh_create(c, g, q, nullptr, c->stream, nullptr, out);
}

namespace {

// The context's rule tables of option q, built and uploaded at the first mesh of that option.
const RuleSet* ruleset(bem2d_ctx* c, bem2d_quad_option q)
{
    std::lock_guard<std::mutex> lock(c->rules_mu);
    for (const RuleSet* r : c->rulesets)
        if (r->N == q.N && r->L == q.L && r->sigma == q.sigma) return r;
    std::vector<double> gx, gw;
    bem2d_internal_gauss01(q.N, gx, gw);
    Rule r1 = adapted1(gx, gw, q.L, q.sigma);
    Rule r2 = adapted23(gx, gw, q.L, q.sigma, false);
    Rule r3 = adapted23(gx, gw, q.L, q.sigma, true);
    std::vector<double> rules;
    for (const Rule* r : {&r1, &r2, &r3}) {
        rules.insert(rules.end(), r->x.begin(), r->x.end());
        rules.insert(rules.end(), r->y.begin(), r->y.end());
        rules.insert(rules.end(), r->w.begin(), r->w.end());
    }
    for (const Rule* r : {&r1, &r2, &r3}) append_abscissa_block(rules, *r);
    std::vector<double> g1d(gx);
    g1d.insert(g1d.end(), gw.begin(), gw.end());
    RuleSet* rs = new RuleSet();
    rs->N = q.N;
    rs->L = q.L;
    rs->sigma = q.sigma;
    rs->m1 = (int)r1.w.size();
    rs->m2 = (int)r2.w.size();
    rs->m3 = (int)r3.w.size();
    double* dr = nullptr;
    double* dg = nullptr;
    if (hipMalloc(&dr, sizeof(double) * (rules.size() + 1)) != hipSuccess ||
        hipMalloc(&dg, sizeof(double) * g1d.size()) != hipSuccess ||
        hipMemcpy(dr, rules.data(), sizeof(double) * rules.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(dg, g1d.data(), sizeof(double) * g1d.size(), hipMemcpyHostToDevice) !=
            hipSuccess) {
        (void)hipFree(dr);
        (void)hipFree(dg);
        delete rs;
        return nullptr;
    }
    rs->d_rules = dr;
    rs->d_g1d = dg;
    c->rulesets.push_back(rs);
    return rs;
}

// Special (coincident / adjacent) entries with the precedence of lib/bem2d_quadrature.h:271-287:
// identical index -> rule 1; else next(i) == index(j) -> rule 3; else prev(i) == index(j) -> 2.
// Elements are looked up by index value through a bucket array over the range of the indices when
// that range is at most a few times n (the usual numbering), else through a hash map.
void special_pairs(const bem2d_geometry* g, std::vector<int64_t>& special)
{
    const int64_t n = g->n;
    int64_t lo = g->index[0], hi = g->index[0];
    for (int64_t e = 1; e < n; ++e) {
        lo = std::min(lo, g->index[e]);
        hi = std::max(hi, g->index[e]);
    }
    const bool dense = hi - lo < 4 * n + 64;
    std::vector<int64_t> start, members;
    std::unordered_map<int64_t, std::vector<int64_t>> pos;
    if (dense) {
        const int64_t range = hi - lo + 1;
        start.assign((size_t)range + 1, 0);
        for (int64_t e = 0; e < n; ++e) ++start[(size_t)(g->index[e] - lo) + 1];
        for (int64_t v = 0; v < range; ++v) start[(size_t)v + 1] += start[(size_t)v];
        members.resize((size_t)n);
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t e = 0; e < n; ++e) members[(size_t)fill[(size_t)(g->index[e] - lo)]++] = e;
    } else {
        pos.reserve((size_t)n * 2);
        for (int64_t e = 0; e < n; ++e) pos[g->index[e]].push_back(e);
    }
    special.clear();
    special.reserve((size_t)n * 9);
    std::vector<int64_t> seen;
    for (int64_t i = 0; i < n; ++i) {
        seen.clear();
        auto add = [&](int64_t idx, int rule) {
            const int64_t* first = nullptr;
            const int64_t* last = nullptr;
            if (dense) {
                if (idx < lo || idx > hi) return;
                first = members.data() + start[(size_t)(idx - lo)];
                last = members.data() + start[(size_t)(idx - lo) + 1];
            } else {
                auto it = pos.find(idx);
                if (it == pos.end()) return;
                first = it->second.data();
                last = first + it->second.size();
            }
            for (const int64_t* p = first; p != last; ++p) {
                const int64_t j = *p;
                bool dup = false;
                for (int64_t s : seen) dup |= (s == j);
                if (dup) continue;
                seen.push_back(j);
                special.push_back(i);
                special.push_back(j);
                special.push_back(rule);
            }
        };
        add(g->index[i], 1);
        add(g->next[i], 3);
        add(g->prev[i], 2);
    }
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

} // namespace

void bem2d_internal_rulesets_release(bem2d_ctx* c)
{
    std::lock_guard<std::mutex> lock(c->rules_mu);
    for (RuleSet* r : c->rulesets) {
        (void)hipFree(r->d_rules);
        (void)hipFree(r->d_g1d);
        delete r;
    }
    c->rulesets.clear();
}

int bem2d_internal_mesh_create(bem2d_ctx* c, const bem2d_geometry* g, bem2d_quad_option q,
                               DevArena* arena, hipStream_t s, HostStager* stager,
                               bem2d_mesh** out)
{
    if (!c || !g || !out) return FAIL(BEM2D_ERR_ARG, "bem2d_mesh_create: null argument");
    if (g->n < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_mesh_create: empty geometry");
    if (q.N < 1 || q.N > kMaxNq)
        return FAIL(BEM2D_ERR_ARG, "QuadOption.N = %d outside [1, %d]", q.N, kMaxNq);
    if (q.L < 0 || q.L > 12) return FAIL(BEM2D_ERR_ARG, "QuadOption.L = %d outside [0, 12]", q.L);
    if (!(q.sigma > 0.0 && q.sigma < 1.0))
        return FAIL(BEM2D_ERR_ARG, "QuadOption.sigma = %g outside (0, 1)", q.sigma);
    if (!g->type || !g->param || !g->index || !g->next || !g->prev)
        return FAIL(BEM2D_ERR_ARG, "bem2d_mesh_create: null geometry array");
    const int64_t n = g->n;
    for (int64_t e = 0; e < n; ++e)
        if (g->type[e] < BEM2D_ELEM_FLAT || g->type[e] > BEM2D_ELEM_INV_ELLIPSE)
            return FAIL(BEM2D_ERR_UNSUPPORTED, "element %lld has unknown type %d", (long long)e,
                        g->type[e]);
    HIPCHK(hipSetDevice(c->device));
    const RuleSet* rs = ruleset(c, q);
    if (!rs) return FAIL(BEM2D_ERR_NOMEM, "bem2d_mesh_create: rule tables");
    std::vector<int64_t> special;
    special_pairs(g, special);

    bem2d_mesh* m = new bem2d_mesh();
    m->ctx = c;
    m->arena = arena;
    m->n = n;
    m->opts = q;
    m->nq = q.N;
    m->m1 = rs->m1;
    m->m2 = rs->m2;
    m->m3 = rs->m3;
    m->nspecial = (int64_t)special.size() / 3;
    m->h_type.assign(g->type, g->type + n);
    m->h_param.assign(g->param, g->param + 8 * n);
    m->d_rules = rs->d_rules;
    m->d_g1d = rs->d_g1d;

    auto cleanup = [&](int code, const char* what) {
        bem2d_mesh_destroy(m);
        return FAIL(code, "bem2d_mesh_create: %s", what);
    };
    // type | param | index, next, prev | special triples: one block, one upload
    const size_t o_param = round256(sizeof(int32_t) * n);
    const size_t o_adj = o_param + round256(sizeof(double) * 8 * n);
    const size_t o_special = o_adj + round256(sizeof(int64_t) * 3 * n);
    const size_t bytes_block = o_special + round256(sizeof(int64_t) * (special.size() + 3));
    const size_t bytes_nodes =
        sizeof(double) * (NODE_PLANES * (size_t)n * q.N + BOUND_PLANES * (size_t)n);
    m->d_block = bem2d_internal_dalloc(arena, bytes_block);
    m->d_nodes = (double*)bem2d_internal_dalloc(arena, bytes_nodes);
    if (!m->d_block || !m->d_nodes) return cleanup(BEM2D_ERR_NOMEM, "device allocation failed");
    char* db = (char*)m->d_block;
    m->d_type = (int32_t*)db;
    m->d_param = (double*)(db + o_param);
    m->d_adj = (int64_t*)(db + o_adj);
    m->d_special = (int64_t*)(db + o_special);
    std::vector<char> local;
    char* h = nullptr;
    if (stager) {
        h = (char*)stager->get(bytes_block);
        if (!h) return cleanup(BEM2D_ERR_NOMEM, "pinned staging");
    } else {
        local.resize(bytes_block);
        h = local.data();
    }
    std::memcpy(h, g->type, sizeof(int32_t) * n);
    std::memcpy(h + o_param, g->param, sizeof(double) * 8 * n);
    std::memcpy(h + o_adj, g->index, sizeof(int64_t) * n);
    std::memcpy(h + o_adj + sizeof(int64_t) * n, g->next, sizeof(int64_t) * n);
    std::memcpy(h + o_adj + 2 * sizeof(int64_t) * n, g->prev, sizeof(int64_t) * n);
    if (!special.empty())
        std::memcpy(h + o_special, special.data(), sizeof(int64_t) * special.size());
    if (hipMemcpyAsync(m->d_block, h, o_special + sizeof(int64_t) * special.size(),
                       hipMemcpyHostToDevice, s) != hipSuccess)
        return cleanup(BEM2D_ERR_DEVICE, "upload failed");
    const int64_t total = n * q.N;
    prep_nodes_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(
        n, q.N, m->d_type, m->d_param, m->d_g1d, m->d_g1d + q.N, m->d_nodes);
    prep_bounds_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, q.N, m->d_nodes,
                                                                   m->bounds());
    if (hipGetLastError() != hipSuccess) return cleanup(BEM2D_ERR_DEVICE, "prep kernel launch");
    // pageable staging goes out of scope on return: wait for the upload
    if (!stager && hipStreamSynchronize(s) != hipSuccess)
        return cleanup(BEM2D_ERR_DEVICE, "prep kernel");
    *out = m;
    return BEM2D_OK;
}

extern "C" void bem2d_mesh_destroy(bem2d_mesh* m)
{
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    if (!m->arena) (void)hipStreamSynchronize(m->ctx->stream);
    DevArena* a = m->arena;
    bem2d_internal_dfree(a, m->d_block);
    bem2d_internal_dfree(a, m->d_nodes);
    delete m;
}

extern "C" int64_t bem2d_mesh_size(const bem2d_mesh* m) { return m ? m->n : -1; }
extern "C" int64_t bem2d_mesh_special_count(const bem2d_mesh* m) { return m ? m->nspecial : -1; }

// ------------------------------------------------------------------------------------------------
// assembly
// ------------------------------------------------------------------------------------------------
static AsmProblem asm_problem(const bem2d_mesh* m, double k, double eta, double mscale,
                              double kscale, double2* A, int64_t lda)
{
    AsmProblem p{};
    p.n = m->n;
    p.nq = m->nq;
    p.m1 = m->m1;
    p.m2 = m->m2;
    p.m3 = m->m3;
    p.nodes = m->d_nodes;
    p.adj = m->d_adj;
    p.special = m->d_special;
    p.nspecial = m->nspecial;
    p.type = m->d_type;
    p.param = m->d_param;
    p.rules = m->d_rules;
    p.g1d = m->d_g1d;
    p.k = k;
    p.eta = eta;
    p.kscale = kscale;
    p.mscale = mscale;
    p.A = A;
    p.lda = lda;
    return p;
}

// Regular and singular-pair kernels for problems[0 .. count) (one kernel kind, one nq), in launches
// of up to kMaxBatch problems each.
template <int KIND>
static int launch_assembly(bem2d_ctx* c, hipStream_t s, const AsmProblem* probs, int count)
{
    const int nq = probs[0].nq;
    for (int c0 = 0; c0 < count; c0 += kMaxBatch) {
        const int cnt = std::min(kMaxBatch, count - c0);
        AsmBatch reg{}, sp{};
        reg.count = sp.count = cnt;
        for (int d = 0; d < cnt; ++d) {
            const AsmProblem& p = probs[c0 + d];
            if (p.nq != nq) return FAIL(BEM2D_ERR_ARG, "batched assembly mixes Gauss orders");
            reg.p[d] = sp.p[d] = p;
            reg.start[d + 1] = reg.start[d] + upper_tile_count(p.n);
            sp.start[d + 1] = sp.start[d] + p.nspecial;
        }
        if (reg.start[cnt] > (int64_t)INT32_MAX || sp.start[cnt] > (int64_t)INT32_MAX)
            return FAIL(BEM2D_ERR_UNSUPPORTED, "assembly grid too large");
        const size_t shm = sizeof(double) * kTileSide * node_stride(nq);
        // The singular-pair kernel writes entries the regular kernel skips: it runs on the second
        // stream beside the regular kernel (forked from and joined back into s).
        const bool fork = sp.start[cnt] > 0;
        if (fork) {
            if (hipEventRecord(c->ev_fork, s) != hipSuccess ||
                hipStreamWaitEvent(c->aux, c->ev_fork, 0) != hipSuccess)
                return FAIL(BEM2D_ERR_DEVICE, "stream fork failed");
            bem2d_internal_prof_begin(c, BEM2D_PROF_ASSEMBLE_SPECIAL, c->aux);
            assemble_special_kernel<KIND><<<(unsigned)sp.start[cnt], 256, 0, c->aux>>>(sp);
            bem2d_internal_prof_end(c, BEM2D_PROF_ASSEMBLE_SPECIAL, c->aux);
        }
        if (reg.start[cnt] > 0) {
            bem2d_internal_prof_begin(c, BEM2D_PROF_ASSEMBLE_REGULAR, s);
            // few tiles (small problems): four wavefronts per tile, each over a quarter of the
            // test nodes, so the launch fills the device and every lane's chain is shorter
            const unsigned grid = (unsigned)reg.start[cnt];
            int split = 1; // the size bucket's choice (autotuner table)
            for (int b = 0; b < c->tune.nasm; ++b)
                if (reg.start[cnt] < c->tune.asm_upper[b]) {
                    split = c->tune.asm_split[b];
                    break;
                }
            if (nq == 16 && split == 4)
                assemble_regular_kernel<KIND, 16, 4><<<grid, 256, shm, s>>>(reg);
            else if (nq == 16 && split == 2)
                assemble_regular_kernel<KIND, 16, 2><<<grid, 128, shm, s>>>(reg);
            else if (nq == 16)
                assemble_regular_kernel<KIND, 16, 1><<<grid, 64, shm, s>>>(reg);
            else if (split == 4)
                assemble_regular_kernel<KIND, 0, 4><<<grid, 256, shm, s>>>(reg);
            else if (split == 2)
                assemble_regular_kernel<KIND, 0, 2><<<grid, 128, shm, s>>>(reg);
            else
                assemble_regular_kernel<KIND, 0, 1><<<grid, 64, shm, s>>>(reg);
            bem2d_internal_prof_end(c, BEM2D_PROF_ASSEMBLE_REGULAR, s);
        }
        if (fork && (hipEventRecord(c->ev_join, c->aux) != hipSuccess ||
                     hipStreamWaitEvent(s, c->ev_join, 0) != hipSuccess))
            return FAIL(BEM2D_ERR_DEVICE, "stream join failed");
    }
    return hipGetLastError() == hipSuccess ? BEM2D_OK : BEM2D_ERR_DEVICE;
}

int bem2d_internal_assemble_many(bem2d_ctx* c, hipStream_t s, int kind,
                                 const bem2d_mesh* const* meshes, const double* k,
                                 const double* eta, double mscale, double kscale,
                                 void* const* A, const int64_t* lda, int count)
{
    if (count <= 0) return BEM2D_OK;
    std::vector<AsmProblem> probs(count);
    for (int d = 0; d < count; ++d) {
        if (!meshes[d] || !A[d] || lda[d] < meshes[d]->n)
            return FAIL(BEM2D_ERR_ARG, "batched assembly: bad problem %d", d);
        if (!(k[d] > 0.0) && kind != BEM2D_KERNEL_NONE && kind != BEM2D_KERNEL_ID)
            return FAIL(BEM2D_ERR_ARG, "wavenumber must be positive (got %g)", k[d]);
        probs[d] = asm_problem(meshes[d], k[d], eta[d], mscale, kscale, (double2*)A[d], lda[d]);
    }
    // group by Gauss order (the kernels are specialised for nq = 16)
    std::stable_sort(probs.begin(), probs.end(),
                     [](const AsmProblem& a, const AsmProblem& b) { return a.nq < b.nq; });
    int rc = BEM2D_OK;
    for (int g0 = 0; g0 < count && rc == BEM2D_OK;) {
        int g1 = g0;
        while (g1 < count && probs[g1].nq == probs[g0].nq) ++g1;
        const AsmProblem* ps = probs.data() + g0;
        const int cnt = g1 - g0;
        switch (kind) {
        case BEM2D_KERNEL_NONE: rc = launch_assembly<BEM2D_KERNEL_NONE>(c, s, ps, cnt); break;
        case BEM2D_KERNEL_SL: rc = launch_assembly<BEM2D_KERNEL_SL>(c, s, ps, cnt); break;
        case BEM2D_KERNEL_DL: rc = launch_assembly<BEM2D_KERNEL_DL>(c, s, ps, cnt); break;
        case BEM2D_KERNEL_CDL: rc = launch_assembly<BEM2D_KERNEL_CDL>(c, s, ps, cnt); break;
        case BEM2D_KERNEL_COMBINED: rc = launch_assembly<BEM2D_KERNEL_COMBINED>(c, s, ps, cnt); break;
        case BEM2D_KERNEL_ID: rc = launch_assembly<BEM2D_KERNEL_ID>(c, s, ps, cnt); break;
        default: return FAIL(BEM2D_ERR_UNSUPPORTED, "unknown kernel kind %d", kind);
        }
        g0 = g1;
    }
    if (rc) return FAIL(rc, "assembly kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    return BEM2D_OK;
}

extern "C" int bem2d_assemble(bem2d_ctx* c, const bem2d_mesh* m, int kind, double k, double eta,
                              double mscale, double kscale, void* A, int64_t lda)
{
    if (!c || !m || !A) return FAIL(BEM2D_ERR_ARG, "bem2d_assemble: null argument");
    if (lda < m->n) return FAIL(BEM2D_ERR_ARG, "bem2d_assemble: lda %lld < n %lld",
                                (long long)lda, (long long)m->n);
    if (!(k > 0.0) && kind != BEM2D_KERNEL_NONE && kind != BEM2D_KERNEL_ID)
        return FAIL(BEM2D_ERR_ARG, "bem2d_assemble: wavenumber must be positive (got %g)", k);
    HIPCHK(hipSetDevice(c->device));
    if (kind < BEM2D_KERNEL_NONE || kind > BEM2D_KERNEL_ID)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "unknown kernel kind %d", kind);
    return bem2d_internal_assemble_many(c, c->stream, kind, &m, &k, &eta, mscale, kscale, &A, &lda,
                                        1);
}

// ------------------------------------------------------------------------------------------------
// right-hand sides, potential, far field
// ------------------------------------------------------------------------------------------------
extern "C" int bem2d_rhs(bem2d_ctx* c, const bem2d_mesh* m, int kind, double k, double eta,
                         const double* dirs, int nrhs, double scale, void* B, int64_t ldb)
{
    if (!c || !m || !B || (!dirs && nrhs > 0)) return FAIL(BEM2D_ERR_ARG, "bem2d_rhs: null argument");
    if (nrhs < 0 || nrhs > 65535 || ldb < m->n) return FAIL(BEM2D_ERR_ARG, "bem2d_rhs: bad sizes");
    if (kind < BEM2D_FIELD_PLANE || kind > BEM2D_FIELD_COMBINED_PLANE)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "unknown field kind %d", kind);
    if (nrhs == 0) return BEM2D_OK;
    HIPCHK(hipSetDevice(c->device));
    const double* d_dirs = nullptr;
    if (int rc = bem2d_internal_stage_dirs(c, dirs, nrhs, &d_dirs)) return rc;
    dim3 grid((unsigned)((m->n + 255) / 256), (unsigned)nrhs);
    rhs_kernel<<<grid, 256, 0, c->stream>>>(m->n, m->nq, m->d_nodes, kind, k, eta, d_dirs, scale,
                                            (double2*)B, ldb);
    HIPCHK(hipGetLastError());
    return BEM2D_OK;
}

// Copies nrhs host directions (x, y) to device scratch in stream order; *out = the device copy.
int bem2d_internal_stage_dirs(bem2d_ctx* c, const double* dirs, int nrhs, const double** out)
{
    const size_t bytes = sizeof(double) * 2 * nrhs;
    double* d_dirs = (double*)bem2d_internal_scratch(c, SCRATCH_SMALL, bytes);
    if (!d_dirs) return FAIL(BEM2D_ERR_NOMEM, "scratch allocation failed");
    // The directions go up through pinned staging memory, so the copy is asynchronous and the call
    // returns before the kernel runs (the caller's array is free again at once).  The staging
    // buffer is refilled only after its previous copy has run.
    if (!c->ev_stage) HIPCHK(hipEventCreateWithFlags(&c->ev_stage, hipEventDisableTiming));
    HIPCHK(hipEventSynchronize(c->ev_stage));
    if (c->stage_bytes < bytes) {
        if (c->stage) HIPCHK(hipHostFree(c->stage));
        c->stage = nullptr;
        c->stage_bytes = 0;
        const size_t want = bytes < 4096 ? 4096 : bytes;
        if (hipHostMalloc((void**)&c->stage, want) != hipSuccess) {
            c->stage = nullptr;
            return FAIL(BEM2D_ERR_NOMEM, "pinned staging allocation failed");
        }
        c->stage_bytes = want;
    }
    std::memcpy(c->stage, dirs, bytes);
    HIPCHK(hipMemcpyAsync(d_dirs, c->stage, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev_stage, c->stream));
    *out = d_dirs;
    return BEM2D_OK;
}

// Right-hand sides with device-resident directions on stream s (no synchronisation).
int bem2d_internal_rhs(const bem2d_mesh* m, int kind, double k, double eta, const double* d_dirs,
                       int nrhs, double scale, void* B, int64_t ldb, hipStream_t s)
{
    if (nrhs <= 0) return BEM2D_OK;
    dim3 grid((unsigned)((m->n + 255) / 256), (unsigned)nrhs);
    rhs_kernel<<<grid, 256, 0, s>>>(m->n, m->nq, m->d_nodes, kind, k, eta, d_dirs, scale,
                                    (double2*)B, ldb);
    return hipGetLastError() == hipSuccess ? BEM2D_OK : FAIL(BEM2D_ERR_DEVICE, "rhs launch failed");
}

// Far field on stream s with a caller-provided signature buffer F (nang x n complex128).
int bem2d_internal_farfield(const bem2d_mesh* m, double k, const void* coef, int64_t ldc,
                            int nrhs, const double* theta, int64_t nang, void* out, void* F,
                            hipStream_t s)
{
    const int64_t tot = nang * m->n;
    if (nang < 1 || nang >= (1 << 24) || tot > (int64_t)UINT32_MAX - 256)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "far field: %lld angles x %lld elements exceed one launch",
                    (long long)nang, (long long)m->n);
    farfield_signature_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(
        m->n, m->nq, m->d_nodes, k, theta, nang, (double2*)F);
    const double pi = 3.14159265358979323846;
    const double mag = -1.0 / std::sqrt(8.0 * pi * k);
    dim3 grid((unsigned)nang, (unsigned)nrhs);
    farfield_combine_kernel<<<grid, 256, 0, s>>>(m->n, (const double2*)F, nang,
                                                 (const double2*)coef, ldc, mag * std::cos(pi / 4),
                                                 mag * std::sin(pi / 4), (double2*)out);
    return hipGetLastError() == hipSuccess ? BEM2D_OK : FAIL(BEM2D_ERR_DEVICE, "far-field launch failed");
}

extern "C" int bem2d_eval_single_layer(bem2d_ctx* c, const bem2d_mesh* m, double k,
                                       const void* coef, const double* pts, int64_t npts, void* out)
{
    if (!c || !m || !coef || (!pts && npts > 0) || (!out && npts > 0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_eval_single_layer: null argument");
    if (!(k > 0.0)) return FAIL(BEM2D_ERR_ARG, "bem2d_eval_single_layer: wavenumber must be positive");
    if (npts <= 0) return BEM2D_OK;
    HIPCHK(hipSetDevice(c->device));
    const bem2d_basis p0{BEM2D_BASIS_CONST, 0, 0.0, 0.0, 0.0};
    BasisSet bs;
    if (int rc = bem2d_internal_make_basis_set(&p0, 1, &bs)) return rc;
    return bem2d_internal_potential(c, m, bs, k, 0.0, coef, pts, npts, out);
}

extern "C" int bem2d_farfield(bem2d_ctx* c, const bem2d_mesh* m, double k, const void* coef,
                              int64_t ldc, int nrhs, const double* theta, int64_t nang, void* out)
{
    if (!c || !m || !coef || !theta || !out) return FAIL(BEM2D_ERR_ARG, "bem2d_farfield: null argument");
    // one 256-thread workgroup per angle: nang * 256 work-items must stay below 2^32
    if (ldc < m->n || nrhs < 1 || nrhs > 65535 || nang < 1 || nang >= (1 << 24) || !(k > 0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_farfield: bad sizes");
    HIPCHK(hipSetDevice(c->device));
    double2* F = (double2*)bem2d_internal_scratch(c, SCRATCH_FARFIELD, sizeof(double2) * nang * m->n);
    if (!F) return FAIL(BEM2D_ERR_NOMEM, "scratch allocation failed");
    const int64_t tot = nang * m->n;
    if (tot > (int64_t)UINT32_MAX - 256)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "far field: %lld angles x %lld elements exceed one launch",
                    (long long)nang, (long long)m->n);
    farfield_signature_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(
        m->n, m->nq, m->d_nodes, k, theta, nang, (double2*)F);
    HIPCHK(hipGetLastError());
    return bem2d_internal_farfield_gemm(c, m->n, k, F, nang, coef, ldc, nrhs, out);
}

// out (nang x nrhs, ld nang) = -(exp(i pi/4) / sqrt(8 pi k)) F coef: the signature matrix F
// (nang x ncols, ld nang) times the coefficients (ncols x nrhs, ld ldc) on rocBLAS zgemm, on the
// context's stream.
int bem2d_internal_farfield_gemm(bem2d_ctx* c, int64_t ncols, double k, const void* F, int64_t nang,
                                 const void* coef, int64_t ldc, int nrhs, void* out)
{
    if (ncols > INT32_MAX || nang > INT32_MAX || ldc > INT32_MAX)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "far field: sizes beyond rocBLAS's int range");
    const double pi = 3.14159265358979323846;
    const double mag = -1.0 / std::sqrt(8.0 * pi * k);
    const rocblas_double_complex alpha(mag * std::cos(pi / 4), mag * std::sin(pi / 4));
    const rocblas_double_complex beta(0.0, 0.0);
    const rocblas_status st = rocblas_zgemm(
        c->blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)nang,
        (rocblas_int)nrhs, (rocblas_int)ncols, &alpha, (const rocblas_double_complex*)F,
        (rocblas_int)nang, (const rocblas_double_complex*)coef, (rocblas_int)ldc, &beta,
        (rocblas_double_complex*)out, (rocblas_int)nang);
    return st == rocblas_status_success ? BEM2D_OK
                                        : FAIL(BEM2D_ERR_DEVICE, "far field zgemm status %d", (int)st);
}

// ------------------------------------------------------------------------------------------------
// plain FP64 solve: rocSOLVER zgetrf + zgetrs
// ------------------------------------------------------------------------------------------------
__global__ void copy_matrix_kernel(int64_t n, const double2* __restrict__ src, int64_t lds,
                                   double2* __restrict__ dst, int64_t ldd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i < n) dst[i + ldd * j] = src[i + lds * j];
}

extern "C" int bem2d_solve_fp64(bem2d_ctx* c, int64_t n, int64_t nrhs, void* A, int64_t lda,
                                void* B, int64_t ldb, int preserve_A)
{
    if (!c || !A || !B) return FAIL(BEM2D_ERR_ARG, "bem2d_solve_fp64: null argument");
    if (n < 1 || nrhs < 1 || lda < n || ldb < n || n > 2000000000 || lda > 2000000000)
        return FAIL(BEM2D_ERR_ARG, "bem2d_solve_fp64: bad sizes");
    HIPCHK(hipSetDevice(c->device));
    double2* F = (double2*)A;
    int64_t ldf = lda;
    if (preserve_A) {
        F = (double2*)bem2d_internal_scratch(c, SCRATCH_FACTOR, sizeof(double2) * n * n);
        if (!F) return FAIL(BEM2D_ERR_NOMEM, "scratch allocation for the factor copy failed");
        ldf = n;
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
        copy_matrix_kernel<<<grid, 256, 0, c->stream>>>(n, (const double2*)A, lda, F, ldf);
        HIPCHK(hipGetLastError());
    }
    int* piv = (int*)bem2d_internal_scratch(c, SCRATCH_PIVOT, sizeof(int) * (n + 1));
    if (!piv) return FAIL(BEM2D_ERR_NOMEM, "scratch allocation failed");
    int* info = piv + n;
    bem2d_internal_prof_begin(c, BEM2D_PROF_LU);
    rocblas_status st = rocsolver_zgetrf(c->blas, (rocblas_int)n, (rocblas_int)n,
                                         (rocblas_double_complex*)F, (rocblas_int)ldf, piv, info);
    bem2d_internal_prof_end(c, BEM2D_PROF_LU);
    if (st != rocblas_status_success) return FAIL(BEM2D_ERR_DEVICE, "rocsolver_zgetrf status %d", (int)st);
    int h_info = 0;
    HIPCHK(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h_info > 0) return FAIL(BEM2D_ERR_SINGULAR, "singular matrix: U(%d,%d) = 0", h_info, h_info);
    st = rocsolver_zgetrs(c->blas, rocblas_operation_none, (rocblas_int)n, (rocblas_int)nrhs,
                          (rocblas_double_complex*)F, (rocblas_int)ldf, piv,
                          (rocblas_double_complex*)B, (rocblas_int)ldb);
    if (st != rocblas_status_success) return FAIL(BEM2D_ERR_DEVICE, "rocsolver_zgetrs status %d", (int)st);
    return BEM2D_OK;
}

// ------------------------------------------------------------------------------------------------
// special functions
// ------------------------------------------------------------------------------------------------
__global__ void hankel01_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a, b, c, d;
    hankel01(x[i], a, b, c, d);
    out[4 * i] = a;
    out[4 * i + 1] = b;
    out[4 * i + 2] = c;
    out[4 * i + 3] = d;
}

extern "C" int bem2d_hankel01(bem2d_ctx* c, const double* x, int64_t n, double* out)
{
    if (!c || (n > 0 && (!x || !out)) || n < 0) return FAIL(BEM2D_ERR_ARG, "bem2d_hankel01: bad argument");
    if (n == 0) return BEM2D_OK;
    HIPCHK(hipSetDevice(c->device));
    hankel01_kernel<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(x, n, out);
    HIPCHK(hipGetLastError());
    return BEM2D_OK;
}

extern "C" int bem2d_hankel01_host(const double* x, int64_t n, double* out)
{
    if (n < 0 || (n > 0 && (!x || !out))) return FAIL(BEM2D_ERR_ARG, "bem2d_hankel01_host: bad argument");
    for (int64_t i = 0; i < n; ++i) hankel01(x[i], out[4 * i], out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);
    return BEM2D_OK;
}
