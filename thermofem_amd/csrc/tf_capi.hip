// C ABI of the library (include/thermofem_hip.h): contexts, memory, the
// kernel-level entry points and the AMG / Krylov handles.  Exceptions never
// cross this boundary; they become negative status codes with a message in
// tf_last_error().
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../include/thermofem/MatrixMarket.hpp"
#include "tf_engine.hpp"
#include "tf_internal.hpp"

namespace tf {
namespace {
thread_local std::string g_error;

// Device copies of host CSR arrays for the host_memory entry points.
struct Staged {
    tf_ctx* c = nullptr;
    int* rp = nullptr;
    int* ci = nullptr;
    double* v = nullptr;
    ~Staged() {
        if (!c) return;
        (void)hipStreamSynchronize(c->stream);
        dev_free(*c, rp);
        dev_free(*c, ci);
        dev_free(*c, v);
    }
};

void stage(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v, Staged& s) {
    s.c = &c;
    s.rp = dev_alloc<int>(c, size_t(n) + 1);
    s.ci = dev_alloc<int>(c, size_t(nnz));
    s.v = dev_alloc<double>(c, size_t(nnz));
    check(hipMemcpyAsync(s.rp, rp, (size_t(n) + 1) * sizeof(int), hipMemcpyHostToDevice, c.stream), "upload rp");
    if (nnz) {
        check(hipMemcpyAsync(s.ci, ci, size_t(nnz) * sizeof(int), hipMemcpyHostToDevice, c.stream), "upload ci");
        check(hipMemcpyAsync(s.v, v, size_t(nnz) * sizeof(double), hipMemcpyHostToDevice, c.stream), "upload v");
    }
}

tf_ctx* make_ctx(int device) {
    int count = 0;
    check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    require(device >= 0 && device < count, "invalid device id");
    DeviceScope scope(device);
    auto c = std::make_unique<tf_ctx>();
    c->device = device;
    check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipMalloc(&c->partials, size_t(kMaxReduceBlocks) * kReduceSlots * sizeof(double)), "hipMalloc partials");
    check(hipMalloc(&c->dscalars, 64 * sizeof(double)), "hipMalloc scalars");
    check(hipHostMalloc(&c->hscalars, 64 * sizeof(double), hipHostMallocDefault), "hipHostMalloc scalars");
    check(hipMalloc(&c->dflags, 64 * sizeof(int)), "hipMalloc flags");
    check(hipHostMalloc(&c->hflags, 64 * sizeof(int), hipHostMallocDefault), "hipHostMalloc flags");
    check(hipDeviceSynchronize(), "context init");
    return c.release();
}

void destroy_ctx(tf_ctx* c) {
    if (!c) return;
    DeviceScope scope(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    c->pool.trim();
    for (int i = 0; i < 32; ++i)
        if (c->ws_ptr[i]) (void)hipFree(c->ws_ptr[i]);
    if (c->scan_tmp) (void)hipFree(c->scan_tmp);
    if (c->partials) (void)hipFree(c->partials);
    if (c->dscalars) (void)hipFree(c->dscalars);
    if (c->hscalars) (void)hipHostFree(c->hscalars);
    if (c->dflags) (void)hipFree(c->dflags);
    if (c->hflags) (void)hipHostFree(c->hflags);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
        (void)hipEventDestroy(c->ev_fork);
        (void)hipEventDestroy(c->ev_join);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

template <class T>
double host_dot(tf_ctx& c, int n, const T* x, const T* y) {
    dot_device<T>(c, n, x, y, c.dscalars, 0);
    check(hipMemcpyAsync(c.hscalars, c.dscalars, sizeof(double), hipMemcpyDeviceToHost, c.stream), "dot download");
    c.sync();
    return c.hscalars[0];
}

}  // namespace

void set_error(const char* msg) { g_error = msg ? msg : ""; }

}  // namespace tf

struct tf_amg {
    tf_ctx* ctx = nullptr;
    tf::AmgEngine* eng = nullptr;
};

struct tf_sell {
    tf_ctx* ctx = nullptr;
    tf::DevSell s;
    double* v64 = nullptr;
    float* v32 = nullptr;
};

struct tf_gnn {
    tf_ctx* ctx = nullptr;
    tf::GnnModel* model = nullptr;
};

struct tf_solver {
    tf_ctx* ctx = nullptr;
    tf::AmgEngine* amg = nullptr;
    tf::SolverEngine* eng = nullptr;
    tf_amg view;
    bool owns = true;  // owns ctx and amg (false: bound to an external tf_amg)
};

using namespace tf;

#define TF_CTX_GUARD(ctx, body)                                      \
    return guard([&] {                                              \
        require((ctx) != nullptr, "null context");                  \
        DeviceScope scope_((ctx)->device);                          \
        body;                                                       \
    })

extern "C" {

const char* tf_last_error(void) { return g_error.c_str(); }

int tf_device_count(int* count) {
    return guard([&] {
        require(count != nullptr, "null count");
        check(hipGetDeviceCount(count), "hipGetDeviceCount");
    });
}

int tf_device_info(int device, char* name, size_t name_len, int* runtime_version) {
    return guard([&] {
        hipDeviceProp_t p{};
        check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
        if (name && name_len) {
            // Some drivers leave the marketing name empty; the ISA name is always set.
            std::strncpy(name, p.name[0] ? p.name : p.gcnArchName, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (runtime_version) check(hipRuntimeGetVersion(runtime_version), "hipRuntimeGetVersion");
    });
}

int tf_sell_create(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, double max_padding,
                   tf_sell** sell) {
    return guard([&] {
        require(ctx != nullptr, "null context");
        DeviceScope scope(ctx->device);
        require(sell != nullptr && n >= 1 && nnz >= 0 && max_padding >= 0, "invalid sliced-ELL arguments");
        *sell = nullptr;
        auto h = std::make_unique<tf_sell>();
        h->ctx = ctx;
        require(sell_build(*ctx, n, nnz, rp, ci, h->s, max_padding),
                "matrix not representable in sliced ELL (row over 255 entries or padding limit)");
        h->v64 = dev_alloc<double>(*ctx, size_t(h->s.stored));
        h->v32 = dev_alloc<float>(*ctx, size_t(h->s.stored));
        sell_values<double, double>(*ctx, h->s, rp, v, h->v64);
        sell_values<double, float>(*ctx, h->s, rp, v, h->v32);
        ctx->sync();
        *sell = h.release();
    });
}

int tf_sell_destroy(tf_sell* h) {
    return guard([&] {
        if (!h) return;
        DeviceScope scope(h->ctx->device);
        h->ctx->sync();
        free_sell(*h->ctx, h->s);
        dev_free(*h->ctx, h->v64);
        dev_free(*h->ctx, h->v32);
        delete h;
    });
}

int tf_sell_info(const tf_sell* h, int* index_bits, int64_t* stored_entries, int* reordered) {
    return guard([&] {
        require(h != nullptr, "null handle");
        if (index_bits) *index_bits = h->s.narrow ? 16 : 32;
        if (stored_entries) *stored_entries = h->s.stored;
        if (reordered) *reordered = h->s.perm != nullptr;
    });
}

#define TF_SELL_ABI(T, SUF, VAL)                                                                                  \
    int tf_sell_spmv_##SUF(tf_ctx* ctx, const tf_sell* h, const T* x, T* y) {                                       \
        TF_CTX_GUARD(ctx, require(h != nullptr && x != y, "SpMV shape or alias mismatch");                           \
                     sell_spmv<T>(*ctx, h->s, h->VAL, x, y));                                                        \
    }                                                                                                                \
    int tf_sell_residual_##SUF(tf_ctx* ctx, const tf_sell* h, const T* b, const T* x, T* r) {                       \
        TF_CTX_GUARD(ctx, require(h != nullptr && r != b && r != x, "Residual shape or alias mismatch");             \
                     sell_residual<T>(*ctx, h->s, h->VAL, b, x, r));                                                 \
    }                                                                                                                \
    int tf_sell_jacobi_##SUF(tf_ctx* ctx, const tf_sell* h, const T* inv, T w, const T* b, const T* x_in, T* x_out) { \
        TF_CTX_GUARD(ctx, require(h != nullptr && x_in != x_out, "fused Jacobi needs distinct input and output iterates"); \
                     sell_jacobi<T>(*ctx, h->s, h->VAL, inv, w, b, x_in, x_out));                                    \
    }
TF_SELL_ABI(double, f64, v64)
TF_SELL_ABI(float, f32, v32)
#undef TF_SELL_ABI

struct tf_stencil {
    tf_ctx* ctx = nullptr;
    tf::DevStencil s;
    double* v64 = nullptr;
    float* v32 = nullptr;
};

int tf_stencil_create(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                      tf_stencil** stencil) {
    return guard([&] {
        require(ctx != nullptr, "null context");
        DeviceScope scope(ctx->device);
        require(stencil != nullptr && n >= 1 && nnz >= 0 && nnz <= INT32_MAX, "invalid stencil arguments");
        *stencil = nullptr;
        require(csr_check(*ctx, n, nnz, rp, ci, v) == 0, "stencil storage needs a canonical CSR matrix");
        auto h = std::make_unique<tf_stencil>();
        h->ctx = ctx;
        require(stencil_build(*ctx, n, nnz, rp, ci, v, h->s, &h->v64),
                "matrix not representable as a stencil (more than 16 distinct column offsets)");
        const size_t count = size_t(h->s.stored);
        h->v32 = dev_alloc<float>(*ctx, count);
        convert<double, float>(*ctx, count, h->v64, h->v32);
        ctx->sync();
        *stencil = h.release();
    });
}

int tf_stencil_destroy(tf_stencil* h) {
    return guard([&] {
        if (!h) return;
        DeviceScope scope(h->ctx->device);
        h->ctx->sync();
        free_stencil(*h->ctx, h->s);
        dev_free(*h->ctx, h->v64);
        dev_free(*h->ctx, h->v32);
        delete h;
    });
}

int tf_stencil_info(const tf_stencil* h, int* terms, int* arrays, int* mirrored) {
    return guard([&] {
        require(h != nullptr, "null handle");
        if (terms) *terms = h->s.terms;
        if (arrays) *arrays = h->s.arrays;
        if (mirrored) *mirrored = h->s.mirrored ? 1 : 0;
    });
}

#define TF_STENCIL_ABI(T, SUF, VAL)                                                                                  \
    int tf_stencil_spmv_##SUF(tf_ctx* ctx, const tf_stencil* h, const T* x, T* y) {                                    \
        TF_CTX_GUARD(ctx, require(h != nullptr && x != y, "SpMV shape or alias mismatch");                              \
                     stencil_spmv<T>(*ctx, h->s, h->VAL, x, y));                                                        \
    }                                                                                                                   \
    int tf_stencil_residual_##SUF(tf_ctx* ctx, const tf_stencil* h, const T* b, const T* x, T* r) {                    \
        TF_CTX_GUARD(ctx, require(h != nullptr && r != b && r != x, "Residual shape or alias mismatch");                \
                     stencil_residual<T>(*ctx, h->s, h->VAL, b, x, r));                                                 \
    }                                                                                                                   \
    int tf_stencil_jacobi_##SUF(tf_ctx* ctx, const tf_stencil* h, const T* inv, T w, const T* b, const T* x_in,        \
                                T* x_out) {                                                                             \
        TF_CTX_GUARD(ctx, require(h != nullptr && x_in != x_out, "fused Jacobi needs distinct input and output iterates"); \
                     stencil_jacobi<T>(*ctx, h->s, h->VAL, inv, w, b, x_in, x_out));                                    \
    }
TF_STENCIL_ABI(double, f64, v64)
TF_STENCIL_ABI(float, f32, v32)
#undef TF_STENCIL_ABI

int tf_read_matrix_market(const char* path, int* rows, int* cols, int64_t* nnz, int* host_rp, int* host_ci,
                          double* host_v) {
    return guard([&] {
        require(path && rows && cols && nnz, "null argument");
        std::ifstream file(path);
        require(bool(file), "Cannot open matrix file");
        const auto a = thermofem::io::readMatrixMarket(file);
        *rows = a.rows_;
        *cols = a.columns_;
        *nnz = int64_t(a.values_.size());
        if (host_rp) std::memcpy(host_rp, a.offsets_.data(), a.offsets_.size() * sizeof(int));
        if (host_ci && !a.indices_.empty()) std::memcpy(host_ci, a.indices_.data(), a.indices_.size() * sizeof(int));
        if (host_v && !a.values_.empty()) std::memcpy(host_v, a.values_.data(), a.values_.size() * sizeof(double));
    });
}

int tf_get_device(int* device) {
    return guard([&] {
        require(device != nullptr, "null device pointer");
        check(hipGetDevice(device), "hipGetDevice");
    });
}

int tf_set_device(int device) {
    return guard([&] { check(hipSetDevice(device), "hipSetDevice"); });
}

int tf_ctx_create(int device, tf_ctx** ctx) {
    return guard([&] {
        require(ctx != nullptr, "null context pointer");
        *ctx = make_ctx(device);
    });
}

int tf_ctx_destroy(tf_ctx* ctx) {
    return guard([&] { destroy_ctx(ctx); });
}

int tf_ctx_synchronize(tf_ctx* ctx) { TF_CTX_GUARD(ctx, ctx->sync()); }
void* tf_ctx_stream(tf_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }
int tf_ctx_device(tf_ctx* ctx) { return ctx ? ctx->device : -1; }

// ---------------------------------------------------------------- events
int tf_event_create(tf_event** ev) {
    return guard([&] {
        require(ev != nullptr, "null event pointer");
        hipEvent_t e = nullptr;
        check(hipEventCreate(&e), "hipEventCreate");
        *ev = reinterpret_cast<tf_event*>(e);
    });
}
int tf_event_destroy(tf_event* ev) {
    return guard([&] {
        if (ev) check(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)), "hipEventDestroy");
    });
}
int tf_event_record(tf_ctx* ctx, tf_event* ev) {
    TF_CTX_GUARD(ctx, check(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), ctx->stream), "hipEventRecord"));
}
int tf_event_elapsed_ms(tf_event* start, tf_event* stop, float* ms) {
    return guard([&] {
        check(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)), "hipEventSynchronize");
        check(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)),
              "hipEventElapsedTime");
    });
}

// ---------------------------------------------------------------- memory
int tf_alloc(int device, int strategy, size_t bytes, void** ptr) {
    return guard([&] {
        require(ptr != nullptr, "null pointer");
        *ptr = nullptr;
        if (!bytes) return;
        DeviceScope scope(device);
        switch (strategy) {
        case TF_MEM_GLOBAL: check(hipMalloc(ptr, bytes), "hipMalloc"); break;
        case TF_MEM_PINNED: check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc"); break;
        case TF_MEM_MAPPED: check(hipHostMalloc(ptr, bytes, hipHostMallocMapped), "hipHostMalloc mapped"); break;
        case TF_MEM_MANAGED: check(hipMallocManaged(ptr, bytes, hipMemAttachGlobal), "hipMallocManaged"); break;
        default: throw Error(TF_E_INVALID, "unknown memory strategy");
        }
    });
}

int tf_free(int strategy, void* ptr) {
    return guard([&] {
        if (!ptr) return;
        if (strategy == TF_MEM_PINNED || strategy == TF_MEM_MAPPED)
            check(hipHostFree(ptr), "hipHostFree");
        else
            check(hipFree(ptr), "hipFree");
    });
}

int tf_device_view(int strategy, void* ptr, void** device_ptr) {
    return guard([&] {
        require(device_ptr != nullptr, "null pointer");
        switch (strategy) {
        case TF_MEM_GLOBAL:
        case TF_MEM_MANAGED: *device_ptr = ptr; break;
        case TF_MEM_MAPPED:
            *device_ptr = nullptr;
            if (ptr) check(hipHostGetDevicePointer(device_ptr, ptr, 0), "hipHostGetDevicePointer");
            break;
        case TF_MEM_PINNED: throw Error(TF_E_INVALID, "Pinned memory has no device view");
        default: throw Error(TF_E_INVALID, "unknown memory strategy");
        }
    });
}

// A plain hipMemcpy may return before a device-to-device copy has finished,
// and the library's streams are non-blocking, so a kernel enqueued next on a
// context stream could read the old contents.  The copy is therefore done on
// the null stream and waited for: the data are in place when this returns.
int tf_memcpy(void* dst, const void* src, size_t bytes) {
    return guard([&] {
        if (!bytes) return;
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, nullptr), "hipMemcpyAsync");
        check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    });
}

int tf_memcpy_async(tf_ctx* ctx, void* dst, const void* src, size_t bytes) {
    TF_CTX_GUARD(ctx, if (bytes) check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream), "hipMemcpyAsync"));
}

int tf_fill_zero(tf_ctx* ctx, void* dst, size_t bytes) { TF_CTX_GUARD(ctx, fill_zero(*ctx, dst, bytes)); }

// ---------------------------------------------------------------- kernels
#define TF_TYPED(SUF, T)                                                                                             \
    int tf_spmv_csr_##SUF(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* x,     \
                          T* y) {                                                                                    \
        TF_CTX_GUARD(ctx, require(x != y || n == 0, "SpMV shape or alias mismatch");                                \
                     spmv<T>(*ctx, n, nnz, rp, ci, v, x, y));                                                        \
    }                                                                                                                \
    int tf_residual_csr_##SUF(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* b, \
                              const T* x, T* r) {                                                                    \
        TF_CTX_GUARD(ctx, require((r != b && r != x) || n == 0, "Residual shape or alias mismatch");                \
                     residual<T>(*ctx, n, nnz, rp, ci, v, b, x, r));                                                 \
    }                                                                                                                \
    int tf_spmm_csr_##SUF(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const T* v, int k,          \
                          const T* X, T* Y) {                                                                        \
        TF_CTX_GUARD(ctx, require(k >= 0 && (X != Y || n == 0), "SpMM shape or alias mismatch");                    \
                     (void)nnz; block_spmm<T>(*ctx, n, rp, ci, v, k, X, Y));                                        \
    }                                                                                                                \
    int tf_jacobi_update_##SUF(tf_ctx* ctx, int n, T w, const T* inv, const T* b, const T* product, T* x) {          \
        TF_CTX_GUARD(ctx, jacobi_update<T>(*ctx, n, w, inv, b, product, x));                                        \
    }                                                                                                                \
    int tf_jacobi_fused_##SUF(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const T* v,             \
                              const T* inv, T w, const T* b, const T* x_in, T* x_out) {                              \
        TF_CTX_GUARD(ctx, require(x_in != x_out, "fused Jacobi needs distinct input and output iterates");          \
                     jacobi_fused<T>(*ctx, n, nnz, rp, ci, v, inv, w, b, x_in, x_out));                             \
    }                                                                                                                \
    int tf_inverse_diagonal_##SUF(tf_ctx* ctx, int n, const int* rp, const int* ci, const T* v, T* inv,            \
                                  int* host_bad) {                                                                   \
        TF_CTX_GUARD(ctx, ctx->clear_flags(1); inverse_diagonal<T>(*ctx, n, rp, ci, v, inv, ctx->dflags);          \
                     ctx->read_flags(1); if (host_bad) *host_bad = ctx->hflags[0]);                                  \
    }                                                                                                                \
    int tf_axpy_##SUF(tf_ctx* ctx, int n, T alpha, const T* x, T* y) {                                               \
        TF_CTX_GUARD(ctx, axpy<T>(*ctx, n, alpha, x, y));                                                           \
    }                                                                                                                \
    int tf_scal_##SUF(tf_ctx* ctx, int n, T alpha, T* x) { TF_CTX_GUARD(ctx, scal<T>(*ctx, n, alpha, x)); }          \
    int tf_dot_##SUF(tf_ctx* ctx, int n, const T* x, const T* y, T* host_result) {                                   \
        TF_CTX_GUARD(ctx, *host_result = T(host_dot<T>(*ctx, n, x, y)));                                            \
    }                                                                                                                \
    int tf_nrm2_##SUF(tf_ctx* ctx, int n, const T* x, T* host_result) {                                             \
        TF_CTX_GUARD(ctx, *host_result = T(std::sqrt(host_dot<T>(*ctx, n, x, x))));                                 \
    }                                                                                                                \
    int tf_restrict_##SUF(tf_ctx* ctx, int nc, const int* offsets, const int* members, const T* fine, T* coarse) {   \
        TF_CTX_GUARD(ctx, require(fine != coarse, "Invalid aggregate transfer vectors");                            \
                     restrict_agg<T>(*ctx, nc, offsets, members, fine, coarse));                                     \
    }                                                                                                                \
    int tf_prolong_##SUF(tf_ctx* ctx, int n, const int* map, const T* coarse, T* fine) {                             \
        TF_CTX_GUARD(ctx, require(fine != coarse, "Invalid aggregate transfer vectors");                            \
                     prolong_agg<T>(*ctx, n, map, coarse, fine, false));                                             \
    }                                                                                                                \
    int tf_coarse_lu_solve_##SUF(tf_ctx* ctx, int n, const double* lu, const int* pivots, const T* b, T* x) {        \
        TF_CTX_GUARD(ctx, coarse_lu_solve<T>(*ctx, n, lu, pivots, b, x));                                           \
    }                                                                                                                \
    int tf_coarse_inv_apply_##SUF(tf_ctx* ctx, int n, const double* inverse, const T* b, T* x) {                     \
        TF_CTX_GUARD(ctx, coarse_inv_apply<T>(*ctx, n, inverse, b, x));                                             \
    }

TF_TYPED(f64, double)
TF_TYPED(f32, float)
#undef TF_TYPED

int tf_convert_f64_f32(tf_ctx* ctx, size_t n, const double* in, float* out) {
    TF_CTX_GUARD(ctx, (convert<double, float>(*ctx, n, in, out)));
}
int tf_convert_f32_f64(tf_ctx* ctx, size_t n, const float* in, double* out) {
    TF_CTX_GUARD(ctx, (convert<float, double>(*ctx, n, in, out)));
}
int tf_convert_i64_i32(tf_ctx* ctx, size_t n, const int64_t* in, int* out) {
    TF_CTX_GUARD(ctx, (convert<int64_t, int>(*ctx, n, in, out)));
}
int tf_convert_i32_i64(tf_ctx* ctx, size_t n, const int* in, int64_t* out) {
    TF_CTX_GUARD(ctx, (convert<int, int64_t>(*ctx, n, in, out)));
}

int tf_copy_f64(tf_ctx* ctx, size_t n, const double* src, double* dst) {
    TF_CTX_GUARD(ctx, (convert<double, double>(*ctx, n, src, dst)));
}

int tf_scan_excl_i32(tf_ctx* ctx, int n, const int* in, int* out) { TF_CTX_GUARD(ctx, scan_excl(*ctx, n, in, out)); }

int tf_scan_excl_i32_noalloc(tf_ctx* ctx, int n, const int* in, int* out) {
    TF_CTX_GUARD(ctx, require(n >= 0, "negative length"); scan_excl_noalloc(*ctx, n, in, out));
}

int tf_csr_check(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int* host_flags) {
    TF_CTX_GUARD(ctx, *host_flags = csr_check(*ctx, n, nnz, rp, ci, v));
}

int tf_csr_validate_structure(tf_ctx* ctx, int rows, int cols, int64_t nnz, const int* rp, const int* ci,
                              int* host_bad) {
    TF_CTX_GUARD(ctx, *host_bad = csr_structure_ok(*ctx, rows, cols, nnz, rp, ci) ? 0 : 1);
}

int tf_csr_canonicalize(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int* out_rp,
                        int** out_ci, double** out_v, int64_t* out_nnz) {
    TF_CTX_GUARD(ctx, {
        DevCsr a = canonicalize(*ctx, n, nnz, rp, ci, v);
        check(hipMemcpyAsync(out_rp, a.rp, (size_t(n) + 1) * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream),
              "canonical offsets");
        ctx->sync();
        dev_free(*ctx, a.rp);
        *out_ci = a.ci;
        *out_v = a.v;
        *out_nnz = a.nnz;
    });
}

namespace {
// Hands a library-made CSR to the caller: offsets into caller memory, columns
// and values as library allocations (released with tf_free).
void hand_out(tf_ctx* ctx, DevCsr& a, int* out_rp, int** out_ci, double** out_v, int64_t* out_nnz) {
    check(hipMemcpyAsync(out_rp, a.rp, (size_t(a.n) + 1) * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream),
          "result offsets");
    ctx->sync();
    dev_free(*ctx, a.rp);
    *out_ci = a.ci;
    *out_v = a.v;
    if (out_nnz) *out_nnz = a.nnz;
}
}  // namespace

int tf_csr_add_f64(tf_ctx* ctx, int rows, int cols, int64_t nnz_a, const int* rp_a, const int* ci_a, const double* v_a,
                   int64_t nnz_b, const int* rp_b, const int* ci_b, const double* v_b, double beta, int* out_rp,
                   int** out_ci, double** out_v, int64_t* out_nnz) {
    TF_CTX_GUARD(ctx, {
        require(rows >= 1 && cols >= 1 && out_rp && out_ci && out_v && out_nnz, "sparse add needs a nonempty shape");
        require(csr_structure_ok(*ctx, rows, cols, nnz_a, rp_a, ci_a) &&
                    csr_structure_ok(*ctx, rows, cols, nnz_b, rp_b, ci_b),
                "invalid CSR operand");
        DevCsr r = csr_add(*ctx, rows, rp_a, ci_a, v_a, rp_b, ci_b, v_b, beta);
        hand_out(ctx, r, out_rp, out_ci, out_v, out_nnz);
    });
}

int tf_csr_transpose_f64(tf_ctx* ctx, int rows, int cols, int64_t nnz, const int* rp, const int* ci, const double* v,
                         int* out_rp, int** out_ci, double** out_v) {
    TF_CTX_GUARD(ctx, {
        require(rows >= 1 && cols >= 1 && out_rp && out_ci && out_v, "sparse transpose needs a nonempty shape");
        require(csr_structure_ok(*ctx, rows, cols, nnz, rp, ci), "invalid CSR operand");
        DevCsr r = csr_transpose(*ctx, rows, cols, nnz, rp, ci, v);
        hand_out(ctx, r, out_rp, out_ci, out_v, nullptr);
    });
}

int tf_csr_multiply_f64(tf_ctx* ctx, int rows_a, int cols_a, int cols_b, int64_t nnz_a, const int* rp_a,
                        const int* ci_a, const double* v_a, int64_t nnz_b, const int* rp_b, const int* ci_b,
                        const double* v_b, int* out_rp, int** out_ci, double** out_v, int64_t* out_nnz) {
    TF_CTX_GUARD(ctx, {
        require(rows_a >= 1 && cols_a >= 1 && cols_b >= 1 && out_rp && out_ci && out_v && out_nnz,
                "sparse product needs nonempty shapes");
        require(csr_structure_ok(*ctx, rows_a, cols_a, nnz_a, rp_a, ci_a) &&
                    csr_structure_ok(*ctx, cols_a, cols_b, nnz_b, rp_b, ci_b),
                "invalid CSR operand");
        DevCsr r = csr_multiply(*ctx, rows_a, rp_a, ci_a, v_a, rp_b, ci_b, v_b);
        hand_out(ctx, r, out_rp, out_ci, out_v, out_nnz);
    });
}

#define TF_DENSE_ABI(T, SUF)                                                                                       \
    int tf_asum_##SUF(tf_ctx* ctx, int n, const T* x, T* host_result) {                                            \
        TF_CTX_GUARD(ctx, require(host_result != nullptr, "null result"); *host_result = T(asum<T>(*ctx, n, x))); \
    }                                                                                                              \
    int tf_dense_transpose_##SUF(tf_ctx* ctx, int rows, int cols, const T* A, T* At) {                             \
        TF_CTX_GUARD(ctx, require(rows >= 0 && cols >= 0 && (A != At || rows * cols == 0), "transpose shape or alias"); \
                     dense_transpose<T>(*ctx, rows, cols, A, At));                                                 \
    }                                                                                                              \
    int tf_gemv_##SUF(tf_ctx* ctx, int rows, int cols, const T* A, const T* x, T* y) {                             \
        TF_CTX_GUARD(ctx, require(rows >= 0 && cols >= 0, "gemv shape"); gemv<T>(*ctx, rows, cols, A, x, y));     \
    }                                                                                                              \
    int tf_gemm_##SUF(tf_ctx* ctx, int m, int n, int k, const T* A, const T* B, T* C) {                            \
        TF_CTX_GUARD(ctx, require(m >= 0 && n >= 0 && k >= 0, "gemm shape"); gemm<T>(*ctx, m, n, k, A, B, C));    \
    }
TF_DENSE_ABI(double, f64)
TF_DENSE_ABI(float, f32)
#undef TF_DENSE_ABI

int tf_agg_validate_rowmax(tf_ctx* ctx, int n, const int* rp, const int* ci, const double* v, double* rowmax,
                           int* host_bad) {
    TF_CTX_GUARD(ctx, *host_bad = validate_rowmax(*ctx, n, rp, ci, v, rowmax) ? 0 : 1);
}

int tf_agg_strength(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, double theta,
                    int** g_rp, int** g_ci, double** g_w, int64_t* g_nnz) {
    TF_CTX_GUARD(ctx, {
        DevCsr g = strength_graph(*ctx, n, nnz, rp, ci, v, theta);
        *g_rp = g.rp;
        *g_ci = g.ci;
        *g_w = g.v;
        *g_nnz = g.nnz;
    });
}

int tf_agg_match_round(tf_ctx* ctx, int n, const int* g_rp, const int* g_ci, const double* g_w, int* mate,
                       int* proposal) {
    TF_CTX_GUARD(ctx, match_round(*ctx, n, g_rp, g_ci, g_w, mate, proposal));
}

int tf_agg_pair(tf_ctx* ctx, int n, const int* g_rp, const int* g_ci, const double* g_w, int rounds, int* map,
                int* host_count) {
    TF_CTX_GUARD(ctx, require(rounds >= 1 && rounds <= 8 && n >= 1, "Invalid matching configuration");
                 *host_count = pair(*ctx, n, g_rp, g_ci, g_w, rounds, map));
}

int tf_agg_contract(tf_ctx* ctx, int n, const int* g_rp, const int* g_ci, const double* g_w, const int* map, int nc,
                    int** c_rp, int** c_ci, double** c_w, int64_t* c_nnz) {
    TF_CTX_GUARD(ctx, {
        DevCsr g = contract_graph(*ctx, n, g_rp, g_ci, g_w, map, nc);
        *c_rp = g.rp;
        *c_ci = g.ci;
        *c_w = g.v;
        *c_nnz = g.nnz;
    });
}

int tf_agg_aggregate(tf_ctx* ctx, int n, const int* g_rp, const int* g_ci, const double* g_w, int rounds, int absorb,
                     int** map, int** offsets, int** members, int* host_coarse_rows, int64_t* host_histogram) {
    TF_CTX_GUARD(ctx, {
        DevCsr g;
        g.n = n;
        g.rp = const_cast<int*>(g_rp);
        g.ci = const_cast<int*>(g_ci);
        g.v = const_cast<double*>(g_w);
        DevAgg t = aggregate(*ctx, g, rounds, absorb != 0);
        *map = t.map;
        *offsets = t.offsets;
        *members = t.members;
        *host_coarse_rows = t.nc;
        if (host_histogram)
            for (int k = 0; k < 4; ++k) host_histogram[k] = t.hist[k];
    });
}

int tf_galerkin(tf_ctx* ctx, int n, const int* rp, const int* ci, const double* v, const int* map, int nc,
                const int* offsets, const int* members, int** c_rp, int** c_ci, double** c_v, int64_t* c_nnz) {
    TF_CTX_GUARD(ctx, {
        DevCsr a;
        a.n = n;
        a.rp = const_cast<int*>(rp);
        a.ci = const_cast<int*>(ci);
        a.v = const_cast<double*>(v);
        a.nnz = -1;  // the members' rows are not known to partition A: staging sized by a read-back
        DevCsr out = galerkin(*ctx, a, map, nc, offsets, members);
        *c_rp = out.rp;
        *c_ci = out.ci;
        *c_v = out.v;
        *c_nnz = out.nnz;
    });
}

int tf_coarse_lu_factor_f64(tf_ctx* ctx, int n, const int* rp, const int* ci, const double* v, double* lu,
                            int* pivots, double* inverse, int* host_singular) {
    TF_CTX_GUARD(ctx, {
        const bool ok = coarse_lu_factor(*ctx, n, rp, ci, v, lu, pivots, inverse);
        ctx->sync();
        *host_singular = ok ? 0 : 1;
    });
}

// ---------------------------------------------------------------- AMG
int tf_amg_create(int device, const tf_amg_options* options, tf_amg** amg) {
    return guard([&] {
        require(amg != nullptr, "null handle pointer");
        const AmgOptionsC o = amg_options_from(options);
        auto h = std::make_unique<tf_amg>();
        h->ctx = make_ctx(device);
        h->eng = new AmgEngine(h->ctx, o);
        *amg = h.release();
    });
}

int tf_amg_destroy(tf_amg* amg) {
    return guard([&] {
        if (!amg) return;
        {
            DeviceScope scope(amg->ctx->device);
            delete amg->eng;
        }
        destroy_ctx(amg->ctx);
        delete amg;
    });
}

int tf_amg_setup(tf_amg* amg, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int host_memory) {
    return guard([&] {
        require(amg != nullptr, "null handle");
        require(n >= 1, "Solver requires a nonempty square matrix");
        DeviceScope scope(amg->ctx->device);
        if (host_memory) {
            Staged s;
            stage(*amg->ctx, n, nnz, rp, ci, v, s);
            amg->eng->setup(n, nnz, s.rp, s.ci, s.v);
            amg->ctx->sync();
        }
        else {
            amg->eng->setup(n, nnz, rp, ci, v);
            amg->ctx->sync();
        }
    });
}

int tf_amg_apply(tf_amg* amg, const double* rhs, double* x, int host_memory) {
    return guard([&] {
        require(amg != nullptr && amg->eng->ready(), "AMG requires setup");
        DeviceScope scope(amg->ctx->device);
        const int n = amg->eng->n();
        if (host_memory) {
            double* d = dev_alloc<double>(*amg->ctx, size_t(n) * 2);
            struct Hold {
                tf_ctx* c;
                double* p;
                ~Hold() {
                    (void)hipStreamSynchronize(c->stream);
                    dev_free(*c, p);
                }
            } hold{amg->ctx, d};
            check(hipMemcpyAsync(d, rhs, size_t(n) * sizeof(double), hipMemcpyHostToDevice, amg->ctx->stream),
                  "upload rhs");
            amg->eng->apply(d, d + n);
            check(hipMemcpyAsync(x, d + n, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, amg->ctx->stream),
                  "download x");
            amg->ctx->sync();
        }
        else {
            amg->eng->apply(rhs, x);
            amg->ctx->sync();
        }
    });
}

int tf_amg_level_count(tf_amg* amg) { return amg && amg->eng ? int(amg->eng->levels_.size()) : 0; }

int tf_amg_level_info(tf_amg* amg, int level, int64_t* rows, int64_t* nnz, int64_t* histogram) {
    return guard([&] {
        require(amg && level >= 0 && level < int(amg->eng->levels_.size()), "level out of range");
        const auto& l = amg->eng->levels_[size_t(level)];
        *rows = l.native.n;
        *nnz = l.native.nnz;
        if (histogram)
            for (int k = 0; k < 4; ++k) histogram[k] = l.hist[k];
    });
}

int tf_amg_level_matrix(tf_amg* amg, int level, int* host_rp, int* host_ci, double* host_v) {
    return guard([&] {
        require(amg && level >= 0 && level < int(amg->eng->levels_.size()), "level out of range");
        DeviceScope scope(amg->ctx->device);
        const auto& a = amg->eng->levels_[size_t(level)].native;
        check(hipMemcpy(host_rp, a.rp, (size_t(a.n) + 1) * sizeof(int), hipMemcpyDeviceToHost), "download rp");
        if (a.nnz) {
            check(hipMemcpy(host_ci, a.ci, size_t(a.nnz) * sizeof(int), hipMemcpyDeviceToHost), "download ci");
            check(hipMemcpy(host_v, a.v, size_t(a.nnz) * sizeof(double), hipMemcpyDeviceToHost), "download v");
        }
    });
}

int tf_amg_level_transfer(tf_amg* amg, int level, int* host_map, int* host_offsets, int* host_members) {
    return guard([&] {
        require(amg && level >= 0 && level + 1 < int(amg->eng->levels_.size()), "level has no transfer");
        DeviceScope scope(amg->ctx->device);
        const auto& t = amg->eng->levels_[size_t(level)].transfer;
        check(hipMemcpy(host_map, t.map, size_t(t.n) * sizeof(int), hipMemcpyDeviceToHost), "download map");
        check(hipMemcpy(host_offsets, t.offsets, (size_t(t.nc) + 1) * sizeof(int), hipMemcpyDeviceToHost),
              "download offsets");
        check(hipMemcpy(host_members, t.members, size_t(t.n) * sizeof(int), hipMemcpyDeviceToHost),
              "download members");
    });
}

const char* tf_amg_stop_reason(tf_amg* amg) { return amg && amg->eng ? amg->eng->stop_.c_str() : ""; }
const char* tf_amg_coarse_solver(tf_amg* amg) { return amg && amg->eng ? amg->eng->coarseSolver_.c_str() : ""; }

int tf_amg_visits(tf_amg* amg, int64_t* visits) {
    return guard([&] {
        require(amg != nullptr, "null handle");
        for (size_t i = 0; i < amg->eng->visits_.size(); ++i) visits[i] = amg->eng->visits_[i];
    });
}

// ---------------------------------------------------------------- solvers
int tf_solver_create(int device, int kind, const tf_krylov_options* options, const tf_amg_options* amg_options,
                     tf_solver** solver) {
    return guard([&] {
        require(solver != nullptr, "null handle pointer");
        KrylovOptionsC o;
        if (options) {
            o.maxit = options->max_iterations;
            o.rtol = options->relative_tolerance;
            o.atol = options->absolute_tolerance;
            o.refresh = options->residual_refresh_interval;
            o.graph = options->use_graph != 0;
        }
        AmgOptionsC ao;
        if (amg_options) ao = amg_options_from(amg_options);
        auto h = std::make_unique<tf_solver>();
        h->ctx = make_ctx(device);
        try {
            DeviceScope scope(device);
            if (amg_options) h->amg = new AmgEngine(h->ctx, ao);
            h->eng = new SolverEngine(h->ctx, kind, o, h->amg);
        }
        catch (...) {
            delete h->amg;
            destroy_ctx(h->ctx);
            throw;
        }
        h->view.ctx = h->ctx;
        h->view.eng = h->amg;
        *solver = h.release();
    });
}

int tf_solver_create_with_amg(int kind, const tf_krylov_options* options, tf_amg* amg, tf_solver** solver) {
    return guard([&] {
        require(solver != nullptr && amg != nullptr, "null handle");
        KrylovOptionsC o;
        if (options) {
            o.maxit = options->max_iterations;
            o.rtol = options->relative_tolerance;
            o.atol = options->absolute_tolerance;
            o.refresh = options->residual_refresh_interval;
            o.graph = options->use_graph != 0;
        }
        auto h = std::make_unique<tf_solver>();
        h->ctx = amg->ctx;
        h->amg = amg->eng;
        h->owns = false;
        DeviceScope scope(h->ctx->device);
        h->eng = new SolverEngine(h->ctx, kind, o, h->amg);
        h->view = *amg;
        *solver = h.release();
    });
}

int tf_solver_destroy(tf_solver* s) {
    return guard([&] {
        if (!s) return;
        {
            DeviceScope scope(s->ctx->device);
            delete s->eng;
            if (s->owns) delete s->amg;
        }
        if (s->owns) destroy_ctx(s->ctx);
        delete s;
    });
}

namespace {
double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    return double(ms) * 1e-3;
}
struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events() {
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventCreate(&b), "hipEventCreate");
    }
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
}  // namespace

int tf_solver_setup(tf_solver* s, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int host_memory) {
    return guard([&] {
        require(s != nullptr, "null handle");
        require(n >= 1, "Solver requires a nonempty square matrix");
        DeviceScope scope(s->ctx->device);
        Staged st;
        if (host_memory) {
            stage(*s->ctx, n, nnz, rp, ci, v, st);
            rp = st.rp;
            ci = st.ci;
            v = st.v;
            s->ctx->sync();
        }
        Events ev;
        check(hipEventRecord(ev.a, s->ctx->stream), "event record");
        s->eng->setup(n, nnz, rp, ci, v);
        check(hipEventRecord(ev.b, s->ctx->stream), "event record");
        check(hipEventSynchronize(ev.b), "event sync");
        s->eng->setupSeconds = elapsed(ev.a, ev.b);
    });
}

int tf_solver_apply(tf_solver* s, const double* b, double* x, int host_memory, tf_solve_result* result) {
    return guard([&] {
        require(s != nullptr && result != nullptr, "null handle");
        DeviceScope scope(s->ctx->device);
        const int n = s->eng->n();
        require(n > 0, "solver requires setup");
        Events ev;
        double* hx = x;
        struct Hold {
            tf_ctx* c;
            double* p;
            ~Hold() {
                (void)hipStreamSynchronize(c->stream);
                dev_free(*c, p);
            }
        } hold{s->ctx, nullptr};
        if (host_memory) {
            double* d = dev_alloc<double>(*s->ctx, size_t(n) * 2);
            hold.p = d;
            check(hipMemcpyAsync(d, b, size_t(n) * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream), "upload b");
            check(hipMemcpyAsync(d + n, x, size_t(n) * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream),
                  "upload x");
            b = d;
            x = d + n;
        }
        check(hipEventRecord(ev.a, s->ctx->stream), "event record");
        *result = s->eng->apply(b, x);
        check(hipEventRecord(ev.b, s->ctx->stream), "event record");
        check(hipEventSynchronize(ev.b), "event sync");
        s->eng->applySeconds = elapsed(ev.a, ev.b);
        if (host_memory) {
            check(hipMemcpyAsync(hx, x, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream),
                  "download x");
            s->ctx->sync();
        }
    });
}

int tf_amg_apply_block(tf_amg* amg, int k, const double* R, double* Z) {
    return guard([&] {
        require(amg != nullptr && amg->eng->ready(), "AMG requires setup");
        require(k >= 1 && R != nullptr && Z != nullptr && static_cast<const void*>(R) != Z,
                "block apply needs k >= 1 and distinct device blocks");
        DeviceScope scope(amg->ctx->device);
        amg->eng->block().apply(k, R, Z);
        amg->ctx->sync();
    });
}

int tf_amg_block_visits(tf_amg* amg, int64_t* visits) {
    return guard([&] {
        require(amg != nullptr && visits != nullptr, "null handle");
        const auto& v = amg->eng->block().visits;
        for (size_t d = 0; d < v.size(); ++d) visits[d] = v[d];
    });
}

int tf_solver_apply_block(tf_solver* s, int k, const double* B, double* X, tf_solve_result* results) {
    return guard([&] {
        require(s != nullptr && results != nullptr, "null handle");
        require(k >= 1 && B != nullptr && X != nullptr && static_cast<const void*>(B) != X,
                "block solve needs k >= 1 and distinct device blocks");
        DeviceScope scope(s->ctx->device);
        Events ev;
        check(hipEventRecord(ev.a, s->ctx->stream), "event record");
        s->eng->apply_block(k, B, X, results);
        check(hipEventRecord(ev.b, s->ctx->stream), "event record");
        check(hipEventSynchronize(ev.b), "event sync");
        s->eng->applySeconds = elapsed(ev.a, ev.b);
    });
}

int tf_solver_apply_segmented(tf_solver* s, int segments, const int* host_offsets, const double* b, double* x,
                              tf_solve_result* results) {
    return guard([&] {
        require(s != nullptr && results != nullptr && host_offsets != nullptr, "null handle");
        require(b != nullptr && x != nullptr && static_cast<const void*>(b) != x,
                "segmented solve needs distinct device vectors");
        DeviceScope scope(s->ctx->device);
        Events ev;
        check(hipEventRecord(ev.a, s->ctx->stream), "event record");
        s->eng->apply_segmented(segments, host_offsets, b, x, results);
        check(hipEventRecord(ev.b, s->ctx->stream), "event record");
        check(hipEventSynchronize(ev.b), "event sync");
        s->eng->applySeconds = elapsed(ev.a, ev.b);
    });
}

int tf_solver_apply_segmented_persistent(tf_solver* s, int segments, const int* host_offsets, const double* b,
                                         double* x, tf_solve_result* results) {
    return guard([&] {
        require(s != nullptr && results != nullptr && host_offsets != nullptr, "null handle");
        require(b != nullptr && x != nullptr && static_cast<const void*>(b) != x,
                "segmented solve needs distinct device vectors");
        DeviceScope scope(s->ctx->device);
        Events ev;
        check(hipEventRecord(ev.a, s->ctx->stream), "event record");
        s->eng->apply_segmented_persistent(segments, host_offsets, b, x, results);
        check(hipEventRecord(ev.b, s->ctx->stream), "event record");
        check(hipEventSynchronize(ev.b), "event sync");
        s->eng->applySeconds = elapsed(ev.a, ev.b);
    });
}

// ------------------------------------------------------- learned affinities

int tf_gnn_param_count(void) { return GnnModel::param_count(); }

int tf_gnn_create(int device, tf_gnn** out) {
    return guard([&] {
        require(out != nullptr, "null handle pointer");
        auto h = std::make_unique<tf_gnn>();
        h->ctx = make_ctx(device);
        DeviceScope scope(device);
        h->model = new GnnModel(h->ctx);
        *out = h.release();
    });
}

int tf_gnn_destroy(tf_gnn* g) {
    return guard([&] {
        if (!g) return;
        {
            DeviceScope scope(g->ctx->device);
            delete g->model;
        }
        destroy_ctx(g->ctx);
        delete g;
    });
}

int tf_gnn_set_params(tf_gnn* g, const float* host, int count) {
    return guard([&] {
        require(g != nullptr && host != nullptr && count == GnnModel::param_count(), "parameter count mismatch");
        DeviceScope scope(g->ctx->device);
        g->model->set_params(host);
    });
}

int tf_gnn_get_params(tf_gnn* g, float* host, int count) {
    return guard([&] {
        require(g != nullptr && host != nullptr && count == GnnModel::param_count(), "parameter count mismatch");
        DeviceScope scope(g->ctx->device);
        g->model->get_params(host);
    });
}

int tf_gnn_infer(tf_gnn* g, int n, int64_t nnz, const int* rp, const int* ci, const double* v, double* weights) {
    return guard([&] {
        require(g != nullptr && weights != nullptr, "null handle");
        DeviceScope scope(g->ctx->device);
        g->model->infer(n, nnz, rp, ci, v, weights);
        g->ctx->sync();
    });
}

int tf_gnn_loss_grad(tf_gnn* g, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                     const float* target, double* loss, float* host_grad) {
    return guard([&] {
        require(g != nullptr && target != nullptr && loss != nullptr, "null handle");
        DeviceScope scope(g->ctx->device);
        *loss = g->model->loss_grad(n, nnz, rp, ci, v, target, host_grad);
    });
}

int tf_gnn_train_step(tf_gnn* g, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                      const float* target, float learning_rate, double* loss) {
    return guard([&] {
        require(g != nullptr && target != nullptr && loss != nullptr, "null handle");
        require(std::isfinite(learning_rate) && learning_rate > 0, "learning rate must be positive");
        DeviceScope scope(g->ctx->device);
        *loss = g->model->train_step(n, nnz, rp, ci, v, target, learning_rate);
    });
}

int tf_affinities(tf_ctx* ctx, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int k, int sweeps,
                  double omega, const double* X0, float* out) {
    TF_CTX_GUARD(ctx, require(X0 != nullptr && out != nullptr, "null pointer");
                 affinities(*ctx, n, nnz, rp, ci, v, k, sweeps, omega, X0, out));
}

int tf_amg_set_affinity_model(tf_amg* amg, tf_gnn* model) {
    return guard([&] {
        require(amg != nullptr, "null handle");
        require(!model || model->ctx->device == amg->ctx->device, "the model must live on the AMG's device");
        amg->eng->set_affinity_model(model ? model->model : nullptr);
    });
}

int tf_amg_learned_info(tf_amg* amg, int* used, double* rho_learned, double* rho_plain) {
    return guard([&] {
        require(amg != nullptr, "null handle");
        if (used) *used = amg->eng->learned_used() ? 1 : 0;
        if (rho_learned) *rho_learned = amg->eng->rho_learned();
        if (rho_plain) *rho_plain = amg->eng->rho_plain();
    });
}

int tf_amg_convergence_estimate(tf_amg* amg, double* rho) {
    return guard([&] {
        require(amg != nullptr && rho != nullptr && amg->eng->ready(), "AMG requires setup");
        DeviceScope scope(amg->ctx->device);
        *rho = amg->eng->convergence_estimate();
    });
}

tf_amg* tf_solver_amg(tf_solver* s) { return s && s->amg ? &s->view : nullptr; }

int tf_solver_timings(tf_solver* s, double* setup_seconds, double* apply_seconds) {
    return guard([&] {
        require(s != nullptr, "null handle");
        if (setup_seconds) *setup_seconds = s->eng->setupSeconds;
        if (apply_seconds) *apply_seconds = s->eng->applySeconds;
    });
}

}  // extern "C"
