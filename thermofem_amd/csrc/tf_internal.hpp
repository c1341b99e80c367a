// Host-side internals shared by the library's translation units: the
// per-device context, error translation and typed launch wrappers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "../../include/thermofem_hip.h"
#include "tf_device.hpp"

namespace tf {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(TF_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void set_error(const char* msg);

// Run f, translating exceptions into status codes for the C ABI.
template <class F>
int guard(F&& f) {
    try {
        f();
        return TF_OK;
    }
    catch (const Error& e) {
        set_error(e.what());
        return e.code;
    }
    catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return TF_E_CAPACITY;
    }
    catch (const std::exception& e) {
        set_error(e.what());
        return TF_E_INVALID;
    }
}

inline void require(bool ok, const char* msg, int code = TF_E_INVALID) {
    if (!ok) throw Error(code, msg);
}

constexpr int kMaxReduceBlocks = 16384;
constexpr int kReduceSlots = 64;

// Scoped current-device switch that restores the caller's device
// (OperationScope, src/LinearAlgebra/Backend.cu:116-126).
struct DeviceScope {
    int previous = -1;
    explicit DeviceScope(int device) {
        check(hipGetDevice(&previous), "hipGetDevice");
        if (previous != device) check(hipSetDevice(device), "hipSetDevice");
    }
    ~DeviceScope() {
        int now = -1;
        if (hipGetDevice(&now) == hipSuccess && now != previous && previous >= 0) (void)hipSetDevice(previous);
    }
};

}  // namespace tf

// Per-context cache of device allocations.  Blocks released by the
// library's own structures (level matrices, transfers, vectors) are kept and
// handed out again for a request of similar size, so repeated setups do not
// pay hipMalloc/hipFree per level.  Blocks are reused only by the context
// that released them; its single stream orders old and new users.
struct DevicePool {
    std::multimap<size_t, void*> free_;
    std::unordered_map<void*, size_t> live_;
    void* get(size_t bytes) {
        bytes = (bytes + 255) / 256 * 256;
        if (bytes == 0) bytes = 256;
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= bytes + bytes / 4 + 65536) {
            void* p = it->second;
            live_[p] = it->first;
            free_.erase(it);
            return p;
        }
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            tf::check(hipMalloc(&p, bytes), "hipMalloc");
        }
        live_[p] = bytes;
        return p;
    }
    void put(void* p) {
        if (!p) return;
        auto it = live_.find(p);
        if (it == live_.end()) {
            (void)hipFree(p);
            return;
        }
        free_.insert({it->second, p});
        live_.erase(it);
    }
    void trim() {
        (void)hipDeviceSynchronize();
        for (auto& kv : free_) (void)hipFree(kv.second);
        free_.clear();
    }
};

struct tf_ctx {
    int device = 0;
    DevicePool pool;
    hipStream_t stream = nullptr;
    double* partials = nullptr;   // per-block partial sums (kMaxReduceBlocks x kReduceSlots slots)
    double* dscalars = nullptr;   // device result scalars
    double* hscalars = nullptr;   // pinned host mirror
    int* dflags = nullptr;        // device integer flags/counters
    int* hflags = nullptr;        // pinned host mirror
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    int spmv_layout = 0;       // 0: LDS-staged row blocks, 1: sub-wave row groups
    void* ws_ptr[32] = {};     // grow-only setup workspace slots
    size_t ws_bytes[32] = {};

    // Side stream for setup work that overlaps the main stream (default
    // priority, created on first use) and the events that fork and join it.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipStream_t side_stream() {
        if (!side) {
            tf::check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreate side");
            tf::check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate fork");
            tf::check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreate join");
        }
        return side;
    }

    void sync() { tf::check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
    // Grow-only scratch for scans.
    int* scan_scratch(size_t ints) {
        const size_t bytes = ints * sizeof(int);
        if (bytes > scan_tmp_bytes) {
            if (scan_tmp) {
                sync();
                tf::check(hipFree(scan_tmp), "hipFree");
            }
            scan_tmp = nullptr;
            tf::check(hipMalloc(&scan_tmp, bytes), "hipMalloc scan scratch");
            scan_tmp_bytes = bytes;
        }
        return static_cast<int*>(scan_tmp);
    }
    // Grow-only device workspace slot (contents undefined on growth).
    void* ws(int slot, size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes > ws_bytes[slot]) {
            if (ws_ptr[slot]) {
                sync();
                tf::check(hipFree(ws_ptr[slot]), "hipFree workspace");
            }
            ws_ptr[slot] = nullptr;
            ws_bytes[slot] = 0;
            const size_t grown = bytes + bytes / 8;
            tf::check(hipMalloc(&ws_ptr[slot], grown), "hipMalloc workspace");
            ws_bytes[slot] = grown;
        }
        return ws_ptr[slot];
    }
    template <class T>
    T* wst(int slot, size_t count) {
        return static_cast<T*>(ws(slot, count * sizeof(T)));
    }
    // Synchronously reads one device int.
    int read_int(const int* p) {
        tf::check(hipMemcpyAsync(hflags, p, sizeof(int), hipMemcpyDeviceToHost, stream), "int download");
        sync();
        return hflags[0];
    }
    // Reads one device int together with device flags [0, slot): the int is
    // copied into flag slot `slot` on the device, then slots [0, slot] come
    // back in one download (flags raised by earlier kernels are checked
    // without a synchronisation of their own).
    int read_int_with_flags(const int* p, int slot) {
        tf::check(hipMemcpyAsync(dflags + slot, p, sizeof(int), hipMemcpyDeviceToDevice, stream), "int stage");
        read_flags(slot + 1);
        return hflags[slot];
    }
    // Reads n ints from device flags after the queued work completes.
    void read_flags(int n) {
        tf::check(hipMemcpyAsync(hflags, dflags, size_t(n) * sizeof(int), hipMemcpyDeviceToHost, stream),
                  "flag download");
        sync();
    }
    void clear_flags(int n) {
        tf::check(hipMemsetAsync(dflags, 0, size_t(n) * sizeof(int), stream), "flag clear");
    }
};

namespace tf {

// ---- typed kernel launchers (defined in the .hip translation units) ----
template <class T>
void spmv(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* x, T* y);
template <class T>
void residual(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* b, const T* x, T* r);
template <class T>
void jacobi_fused(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const T* inv, T w,
                  const T* b, const T* xin, T* xout);
template <class T>
void jacobi_update(tf_ctx& c, int n, T w, const T* inv, const T* b, const T* product, T* x);
// Y = A X for row-major n x k blocks (tf_block.hip); every column's row
// sums are the single-vector SpMV's (rounded products in stored order).
template <class T>
void block_spmm(tf_ctx& c, int n, const int* rp, const int* ci, const T* v, int k, const T* X, T* Y);
// One weighted-Jacobi sweep on every column: Y = X + (w inv) (B - A X).
template <class T>
void block_jacobi(tf_ctx& c, int n, const int* rp, const int* ci, const T* v, const T* inv, T w, int k, const T* B,
                  const T* X, T* Y);

// Learned edge-affinity network (tf_gnn.hip).  Matrices must have a
// symmetric pattern; weights are per stored entry (0 on the diagonal).
class GnnModel {
public:
    explicit GnnModel(tf_ctx* ctx);
    ~GnnModel();
    GnnModel(const GnnModel&) = delete;
    GnnModel& operator=(const GnnModel&) = delete;
    static int param_count();
    void set_params(const float* host);  // FP32 master weights; resets Adam
    void get_params(float* host);
    // out[nnz] (device): the network's weight of every stored entry
    void infer(int n, int64_t nnz, const int* rp, const int* ci, const double* v, double* out);
    // mean squared error against target[nnz] (device, per stored entry) and its gradient
    double loss_grad(int n, int64_t nnz, const int* rp, const int* ci, const double* v, const float* target,
                     float* hostGrad);
    double train_step(int n, int64_t nnz, const int* rp, const int* ci, const double* v, const float* target,
                      float lr);
    tf_ctx* context() const;

private:
    struct Impl;
    Impl* p_;
};
// Affinities of the stored entries after `sweeps` weighted-Jacobi sweeps on
// A X = 0 from the n x k block X0 (device): out[nnz] (device).
void affinities(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int k, int sweeps,
                double omega, const double* X0, float* out);
template <class T>
void axpy(tf_ctx& c, int n, T alpha, const T* x, T* y);
template <class T>
void scal(tf_ctx& c, int n, T alpha, T* x);
// Sum of x*y into the device scalar dst (slot selects the partial-sum storage).
template <class T>
void dot_device(tf_ctx& c, int n, const T* x, const T* y, double* dst, int slot);
template <class A, class B>
void convert(tf_ctx& c, size_t n, const A* in, B* out);
template <class T>
void restrict_agg(tf_ctx& c, int nc, const int* offsets, const int* members, const T* fine, T* coarse,
                  const T* inv = nullptr, T w = T(0), T* xc = nullptr);
template <class T>
void residual_restrict(tf_ctx& c, int nc, const int* offsets, const int* members, const int* rp, const int* ci,
                       const T* v, const T* b, const T* x, T* coarse, const T* inv, T w, T* xc);
template <class T>
void prolong_jacobi(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const T* v, const int* map,
                    const T* xc, const T* x, bool zero, T* xout, const T* b, const T* inv, T w);
template <class T>
void enter_level(tf_ctx& c, int n, const double* rhs, T* b, const T* inv, T w, T* x);
template <class T>
void prolong_agg(tf_ctx& c, int n, const int* map, const T* coarse, T* fine, bool set = false);
template <class T>
void inverse_diagonal(tf_ctx& c, int n, const int* rp, const int* ci, const T* v, T* inv, int* dbad);
void fill_zero(tf_ctx& c, void* dst, size_t bytes);
void scan_excl(tf_ctx& c, int n, const int* in, int* out);
// The same scan with its block totals in a fixed device array: no allocation,
// n <= 2^28 (tf_scan_coarse.hip).
void scan_excl_noalloc(tf_ctx& c, int n, const int* in, int* out);
template <class T>
void coarse_inv_apply(tf_ctx& c, int n, const double* inverse, const T* b, T* x);
template <class T>
void coarse_lu_solve(tf_ctx& c, int n, const double* lu, const int* piv, const T* b, T* x);
bool coarse_lu_factor(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* lu, int* piv,
                      double* inverse);
// coarse_lu_factor queued on the context's side stream behind the work
// queued so far on its main stream, so later main-stream work overlaps it;
// coarse_lu_join makes the main stream wait for it and returns whether the
// factorisation succeeded.  Nothing the factorisation reads or writes may be
// freed or changed in between.
void coarse_lu_launch(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* lu, int* piv,
                      double* inverse);
bool coarse_lu_join(tf_ctx& c);

// ---- aggregation setup (tf_agg.hip) ----
// Device CSR owned by the caller (hipMalloc'd arrays, released by free_csr).
struct DevCsr {
    int n = 0;
    int64_t nnz = 0;
    int* rp = nullptr;
    int* ci = nullptr;
    double* v = nullptr;
};
void free_csr(tf_ctx& c, DevCsr& a);
// Aggregation transfer (AggregateTransfer, GpuAggregation.cuh:13-17).
struct DevAgg {
    int n = 0, nc = 0;
    int* map = nullptr;
    int* offsets = nullptr;
    int* members = nullptr;
    int64_t hist[4] = {0, 0, 0, 0};
};
void free_agg(tf_ctx& c, DevAgg& t);
// Sliced-ELL copy of a CSR structure (tf_sell.hip); values live in separate
// arrays of the element type the kernels run in.
constexpr int kSellWsSlot = 28;
struct DevSell {
    int n = 0, slices = 0;
    int64_t stored = 0;        // entries including padding
    bool narrow = false;       // 16-bit column offsets from the row
    int* base = nullptr;
    unsigned char* len = nullptr;
    void* col = nullptr;
    int* perm = nullptr;       // row at each position, nullptr when rows keep their order
    int lanes = 1;             // lanes per row of the product kernels (1, 2 or 4; same results)
    bool packed = false;       // entries in groups of 4 per lane (sell_slot)
    template <class C>
    SellView<C> view() const {
        return SellView<C>{n, slices, packed ? 1 : 0, base, len, static_cast<const C*>(col), perm};
    }
    explicit operator bool() const { return base != nullptr; }
};
// Builds the structure; false (nothing kept) if a row has more than 255
// entries or padding would exceed maxPadding * nnz.
// maxPackedPadding (< 0: maxPadding) bounds the padding of the packed
// layout, which matrices of 10+ entries a row on average take first.
bool sell_build(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, DevSell& s, double maxPadding,
                double maxPackedPadding = -1);
// The same with the layout given (pack: groups of 4 entries per lane).
bool sell_build(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, DevSell& s, double maxPadding,
                bool pack);
template <class A, class T>
void sell_values(tf_ctx& c, const DevSell& s, const int* rp, const A* v, T* out);
void free_sell(tf_ctx& c, DevSell& s);
template <class T>
void sell_spmv(tf_ctx& c, const DevSell& s, const T* v, const T* x, T* y);
template <class T>
void sell_residual(tf_ctx& c, const DevSell& s, const T* v, const T* b, const T* x, T* r);
template <class T>
void sell_jacobi(tf_ctx& c, const DevSell& s, const T* v, const T* inv, T w, const T* b, const T* xin, T* xout);

// Structured-stencil copy of a CSR operator (tf_stencil.hip): a 16-bit term
// mask per row and one dense value array per stored offset, no column
// indices; the `arrays` value arrays, interleaved by 64-row tiles (`stored`
// elements in all), live in separate buffers of the element type the
// kernels run in.
struct DevStencil {
    int n = 0, terms = 0, arrays = 0;
    int64_t stored = 0;
    bool mirrored = false;  // symmetric: the negative offsets read the positive ones' arrays
    unsigned short* mask = nullptr;
    int off[kStencilTerms] = {}, shift[kStencilTerms] = {}, src[kStencilTerms] = {};
    StencilView view() const {
        StencilView v{n, 64 * arrays, mask, {}, {}, {}};
        for (int t = 0; t < kStencilTerms; ++t) {
            v.off[t] = off[t];
            v.shift[t] = shift[t];
            v.src[t] = src[t];
        }
        return v;
    }
    // Row-sum variant of the product kernels: 0 the mirrored 7-point stencil
    // (stencil7_row_sum), else 1 or 2 chunks of 8 terms.
    int chunks() const {
        const bool seven = terms == 7 && mirrored && off[2] == -1 && off[3] == 0 && off[4] == 1 &&
                           off[1] == -off[5] && off[0] == -off[6] && off[5] > 1 && off[6] > off[5] &&
                           arrays == 4 && src[3] == 0 && src[4] == 64 && src[5] == 128 && src[6] == 192 &&
                           src[2] == 64 && src[1] == 128 && src[0] == 192 && shift[2] == -1 &&
                           shift[1] == off[1] && shift[0] == off[0] && shift[3] == 0 && shift[4] == 0;
        return seven ? 0 : (terms > 8 ? 2 : 1);
    }
    explicit operator bool() const { return mask != nullptr; }
};
// Builds the storage when every row's columns come from one list of at most
// kStencilTerms offsets (found on a sample of rows, then checked on every
// entry) and fills its FP64 values (*vals64, s.stored doubles); false,
// with nothing kept, otherwise.
bool stencil_build(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v, DevStencil& s,
                   double** vals64);
void free_stencil(tf_ctx& c, DevStencil& s);
// to = from with a mask array of its own (the values are the caller's).
void stencil_copy(tf_ctx& c, const DevStencil& from, DevStencil& to);
template <class T>
void stencil_spmv(tf_ctx& c, const DevStencil& s, const T* v, const T* x, T* y);
template <class T>
void stencil_residual(tf_ctx& c, const DevStencil& s, const T* v, const T* b, const T* x, T* r);
template <class T>
void stencil_jacobi(tf_ctx& c, const DevStencil& s, const T* v, const T* inv, T w, const T* b, const T* xin,
                    T* xout);
template <class T>
T* dev_alloc(tf_ctx& c, size_t count) {
    return static_cast<T*>(c.pool.get((count ? count : 1) * sizeof(T)));
}
inline void dev_free(tf_ctx& c, void* p) { c.pool.put(p); }
int csr_check(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v);
bool csr_structure_ok(tf_ctx& c, int rows, int cols, int64_t nnz, const int* rp, const int* ci);
DevCsr canonicalize(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v);
// Rows of any column range sorted by column, duplicates summed in stored order.
DevCsr merge_csr_rows(tf_ctx& c, int rows, int64_t nnz, const int* rp, const int* ci, const double* v);
// Level-1 API of Operations.cuh (tf_dense.hip).
template <class T>
double asum(tf_ctx& c, int n, const T* x);
template <class T>
void dense_transpose(tf_ctx& c, int rows, int cols, const T* A, T* At);
template <class T>
void gemv(tf_ctx& c, int rows, int cols, const T* A, const T* x, T* y);
template <class T>
void gemm(tf_ctx& c, int m, int n, int k, const T* A, const T* B, T* C);
DevCsr csr_add(tf_ctx& c, int rows, const int* ra, const int* ca, const double* va, const int* rb, const int* cb,
               const double* vb, double beta);
DevCsr csr_transpose(tf_ctx& c, int rows, int cols, int64_t nnz, const int* rp, const int* ci, const double* v);
DevCsr csr_multiply(tf_ctx& c, int rows, const int* ra, const int* ca, const double* va, const int* rb, const int* cb,
                    const double* vb);
DevCsr copy_csr(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v);
bool validate_rowmax(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* rowmax);
// The same checks without reading them back: *flag (zeroed by the caller)
// becomes nonzero when a row fails.
void validate_rowmax_async(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* rowmax,
                           int* flag);
// validate_rowmax_async fused with the inverse diagonal of a canonical matrix
// in precision T (flagInv raised as inverse_diagonal raises its flag).
template <class T>
void validate_inverse_async(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v,
                            double* rowmax, T* inv, int* flag, int* flagInv);
// symmetric: -1 unknown (checked here), 0/1 known pattern symmetry;
// valueSymmetric: the values are known to be symmetric bit for bit (then the
// pattern is too, and a_ji is read as a_ij).
// prevalidated: the caller ran validate_rowmax with rowmax_workspace(c, n).
// With deferred = k > 0 the graph's size comes back through
// read_int_with_flags(..., k): flag slots 1..k-1 (slot 0 is the graph's own
// symmetry check) are in c.hflags when it returns.
DevCsr strength_graph(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v, double theta,
                      int symmetric = -1, bool prevalidated = false, int* symmetricOut = nullptr,
                      const double* learned = nullptr, int deferred = 0, bool valueSymmetric = false);
double* rowmax_workspace(tf_ctx& c, int n);
void match_round(tf_ctx& c, int n, const int* rp, const int* ci, const double* w, int* mate, int* proposal);
// nnz < 0: read from rp[n].  With sizes / slots (capacity n + 1 and 4n + 1)
// the member table of the pairs is written too: sizes[a] and the sorted
// members in slots[4a..], sizes[count] = 0; with inSizes / inSlots (the
// table of a previous pairing whose aggregates are this graph's vertices,
// at most two members each) the members are the union of the paired
// aggregates' members.
int pair(tf_ctx& c, int n, const int* rp, const int* ci, const double* w, int rounds, int* map, int64_t nnz = -1,
         const int* inSizes = nullptr, const int* inSlots = nullptr, int* sizes = nullptr, int* slots = nullptr);
DevCsr contract_graph(tf_ctx& c, int n, const int* rp, const int* ci, const double* w, const int* map, int nc,
                      int64_t nnz = -1);
// With hostHist (pinned, 4 ints) the aggregate-size histogram is copied
// there asynchronously and DevAgg::hist is left zero: it is valid once the
// stream has drained.
DevAgg aggregate(tf_ctx& c, const DevCsr& g, int rounds, bool absorb, int* hostHist = nullptr);
DevCsr galerkin(tf_ctx& c, const DevCsr& a, const int* map, int nc, const int* offsets, const int* members);

}  // namespace tf
