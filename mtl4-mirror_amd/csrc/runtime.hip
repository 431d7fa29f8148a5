// Library state, buffer pool, vectors and matrices of the C ABI.
#include "m4g_internal.hpp"

#include <algorithm>
#include <chrono>

namespace m4g {

static thread_local int tl_device = 0;
static thread_local hipStream_t tl_stream = nullptr;   // user stream (nullptr = library stream)

static std::mutex g_states_mutex;
static std::map<int, std::unique_ptr<DeviceState>> g_states;

// Scratch of m4g_dot_noalloc, the entry point that allocates nothing: the
// reduction partials (at most kReduceBlocks, one per workgroup of the first
// stage) and, after them, the result slot.  One copy per GPU, held until the
// process ends; every m4g_dot_noalloc call on a GPU shares it, so calls that
// overlap on one GPU race.  In host code `scratch` names a host-side stand-in:
// only the address hipGetSymbolAddress returns for it (kept per device in
// DeviceState::noalloc_scratch) may reach a kernel or a copy.
constexpr int kNoAllocResult = kReduceBlocks;          // index of the result slot
constexpr int kNoAllocSlots = kReduceBlocks + 32;      // 8448 bytes
static __device__ __align__(256) double scratch[kNoAllocSlots];

DeviceState& device_state()
{
    std::lock_guard<std::mutex> lock(g_states_mutex);
    auto it = g_states.find(tl_device);
    if (it != g_states.end()) {
        M4G_HIP(hipSetDevice(tl_device));
        return *it->second;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw Error(M4G_ERR_NO_DEVICE, "no HIP device visible");
    M4G_REQUIRE(tl_device >= 0 && tl_device < count, M4G_ERR_INVALID, "device index out of range");
    M4G_HIP(hipSetDevice(tl_device));
    auto st = std::make_unique<DeviceState>();
    st->device = tl_device;
    M4G_HIP(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    M4G_HIP(hipGetDeviceProperties(&prop, tl_device));
    st->num_cus = prop.multiProcessorCount;
    M4G_HIP(hipHostMalloc(reinterpret_cast<void**>(&st->pinned), 64 * sizeof(double), hipHostMallocDefault));
    // this device's copy of the no-allocation scratch (the device is current)
    void* sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !sp)
        throw Error(M4G_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se));
    st->noalloc_scratch = static_cast<double*>(sp);
    DeviceState& ref = *st;
    g_states[tl_device] = std::move(st);
    return ref;
}

DeviceState* existing_device_state()
{
    std::lock_guard<std::mutex> lock(g_states_mutex);
    auto it = g_states.find(tl_device);
    if (it == g_states.end()) return nullptr;
    M4G_HIP(hipSetDevice(tl_device));
    return it->second.get();
}

hipStream_t current_stream()
{
    DeviceState& st = device_state();
    return tl_stream ? tl_stream : st.stream;
}

void* pool_alloc(size_t bytes)
{
    DeviceState& st = device_state();
    {
        std::lock_guard<std::mutex> lock(st.pool_mutex);
        auto it = st.pool.find(bytes);
        if (it != st.pool.end()) {
            void* p = it->second;
            st.pool.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        // release cached buffers and retry once
        {
            std::lock_guard<std::mutex> lock(st.pool_mutex);
            M4G_HIP(hipDeviceSynchronize());
            for (auto& kv : st.pool) (void)hipFree(kv.second);
            st.pool.clear();
        }
        M4G_HIP(hipMalloc(&p, bytes));
    }
    return p;
}

void pool_free(void* p, size_t bytes)
{
    if (!p) return;
    DeviceState& st = device_state();
    std::lock_guard<std::mutex> lock(st.pool_mutex);
    st.pool.emplace(bytes, p);
}

double read_scalar(const double* d)
{
    hipStream_t s = current_stream();
    Mailbox mb(s);
    const unsigned q = mb.arm();
    launch_post(d, 1, mb.box(), q, s);
    return mb.wait(q)[0];
}

Mailbox::Mailbox(hipStream_t s) : s_(s)
{
    DeviceState& st = device_state();
    {
        std::lock_guard<std::mutex> lock(st.mail_mutex);
        if (!st.mailboxes.empty()) {
            box_ = st.mailboxes.back().first;
            armed_ = seen_ = st.mailboxes.back().second;
            st.mailboxes.pop_back();
        }
    }
    if (!box_) {
        M4G_HIP(hipHostMalloc(reinterpret_cast<void**>(&box_), sizeof(HostPost), hipHostMallocCoherent));
        std::memset(static_cast<void*>(box_), 0, sizeof(HostPost));
    }
}

Mailbox::~Mailbox()
{
    // a post still in flight (a solve left by an exception) must land before
    // the mailbox serves another solve
    if (seen_ != armed_ && hipStreamSynchronize(s_) != hipSuccess) return;   // a failed stream: keep it out of the pool
    DeviceState& st = device_state();
    std::lock_guard<std::mutex> lock(st.mail_mutex);
    st.mailboxes.emplace_back(box_, armed_);
}

const double* Mailbox::wait(unsigned seq)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned k = 1;; k++) {
        if (__atomic_load_n(&box_->seq, __ATOMIC_ACQUIRE) == seq) {
            seen_ = seq;
            return box_->val;
        }
        if ((k & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
        __builtin_ia32_pause();
    }
    // a long kernel ahead of the post, or a failed one: the stream's own wait
    M4G_HIP(hipStreamSynchronize(s_));
    M4G_REQUIRE(__atomic_load_n(&box_->seq, __ATOMIC_ACQUIRE) == seq, M4G_ERR_HIP,
                "a posted scalar did not arrive although its stream completed");
    seen_ = seq;
    return box_->val;
}

// The pattern of trans(A) (rows of A's columns, entries in A's row order, so
// each row of the transpose sums in the order of the transposed traversal),
// uploaded, and pos[d] = the position in A of the transpose's entry d.
std::unique_ptr<m4g_matrix_s> transpose_pattern(const m4g_matrix_s& A, DevBuf<int64_t>& pos_dev)
{
    auto T = std::make_unique<m4g_matrix_s>();
    const int64_t m = A.nrows, n = A.ncols, nnz = A.nnz;
    T->nrows = n;
    T->ncols = m;
    T->nnz = nnz;
    T->h_starts.assign(n + 1, 0);
    T->h_indices.assign(nnz, 0);
    std::vector<int64_t> pos(nnz);
    for (int64_t p = 0; p < nnz; p++) T->h_starts[A.h_indices[p] + 1]++;
    for (int64_t r = 0; r < n; r++) T->h_starts[r + 1] += T->h_starts[r];
    std::vector<int64_t> fill(T->h_starts.begin(), T->h_starts.end() - 1);
    for (int64_t r = 0; r < m; r++)
        for (int64_t p = A.h_starts[r]; p < A.h_starts[r + 1]; p++) {
            int64_t d = fill[A.h_indices[p]]++;
            T->h_indices[d] = r;
            pos[d] = p;
        }
    std::vector<int32_t> s32(T->h_starts.begin(), T->h_starts.end()), i32(T->h_indices.begin(), T->h_indices.end());
    T->starts.reset(n + 1);
    T->indices.reset(std::max<int64_t>(nnz, 1));
    T->values.reset(std::max<int64_t>(nnz, 1));
    T->starts.upload(s32.data(), n + 1);
    T->indices.upload(i32.data(), nnz);
    pos_dev.reset(std::max<int64_t>(nnz, 1));
    pos_dev.upload(pos.data(), nnz);
    return T;
}

const m4g_matrix_s& transpose_of(m4g_matrix_s& A, hipStream_t s)
{
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    const bool first = !A.tr;
    if (first) A.tr = transpose_pattern(A, A.tr_pos);
    m4g_matrix_s& T = *A.tr;
    launch_gather_values(A.nnz, A.tr_pos.ptr, A.values.ptr, T.values.ptr, s);
    if (first) sell_build(T);   // lays out the values just gathered
    else sell_refresh_values(T);
    return T;
}

ReduceWs& reduce_ws()
{
    static std::map<int, ReduceWs> ws;
    static std::mutex mu;
    DeviceState& st = device_state();
    std::lock_guard<std::mutex> lock(mu);
    ReduceWs& w = ws[st.device];
    if (!w.partials) {
        M4G_HIP(hipMalloc(&w.partials, sizeof(double) * (kReduceBlocks * 16 + 64)));
        M4G_HIP(hipMalloc(&w.counters, sizeof(unsigned) * kCounters));
        M4G_HIP(hipMemsetAsync(w.counters, 0, sizeof(unsigned) * kCounters, current_stream()));
        M4G_HIP(hipStreamSynchronize(current_stream()));
    }
    return w;
}

// ---------------------------------------------------------------------------
// host-side CSR ingest
// ---------------------------------------------------------------------------
template <typename T>
static void widen(const void* src, int64_t count, std::vector<int64_t>& dst)
{
    const T* s = static_cast<const T*>(src);
    dst.resize(count);
    for (int64_t i = 0; i < count; i++) dst[i] = static_cast<int64_t>(s[i]);
}

static void widen_index(const void* src, int64_t count, int type, std::vector<int64_t>& dst)
{
    switch (type) {
    case M4G_IDX_I32: widen<int32_t>(src, count, dst); break;
    case M4G_IDX_I64: widen<int64_t>(src, count, dst); break;
    case M4G_IDX_U32: widen<uint32_t>(src, count, dst); break;
    default: throw Error(M4G_ERR_INVALID, "unknown index type");
    }
}

}  // namespace m4g

using namespace m4g;

extern "C" {

const char* m4g_version(void)
{
    return "mtl4gpu 0.1 gfx950 (fp-contract=off, packed-fp32-ops disabled)";
}

int m4g_set_device(int device)
{
    return guard([&] {
        tl_device = device;
        device_state();
        return M4G_OK;
    });
}

int m4g_set_stream(void* s)
{
    return guard([&] {
        tl_stream = static_cast<hipStream_t>(s);
        return M4G_OK;
    });
}

void* m4g_get_stream(void)
{
    try {
        return current_stream();
    } catch (...) {
        return nullptr;
    }
}

int m4g_synchronize(void)
{
    return guard([&] {
        M4G_HIP(hipStreamSynchronize(current_stream()));
        return M4G_OK;
    });
}

// ------------------------------- vectors -----------------------------------
int m4g_vector_create(m4g_vector* v, int64_t n, int type)
{
    return guard([&] {
        M4G_REQUIRE(v && n >= 0 && (type == M4G_F64 || type == M4G_F32), M4G_ERR_INVALID,
                    "m4g_vector_create: bad argument");
        auto h = std::make_unique<m4g_vector_s>();
        h->n = n;
        h->type = type;
        h->owned = true;
        h->ptr = n ? pool_alloc(h->bytes()) : nullptr;
        *v = h.release();
        return M4G_OK;
    });
}

int m4g_vector_wrap(m4g_vector* v, void* dptr, int64_t n, int type)
{
    return guard([&] {
        M4G_REQUIRE(v && n >= 0 && (dptr || n == 0) && (type == M4G_F64 || type == M4G_F32),
                    M4G_ERR_INVALID, "m4g_vector_wrap: bad argument");
        auto h = std::make_unique<m4g_vector_s>();
        h->n = n;
        h->type = type;
        h->owned = false;
        h->ptr = dptr;
        *v = h.release();
        return M4G_OK;
    });
}

int m4g_vector_destroy(m4g_vector v)
{
    return guard([&] {
        if (!v) return M4G_OK;
        if (v->owned && v->ptr) {
            // queued work may still use the buffer: pool reuse is stream ordered
            pool_free(v->ptr, v->bytes());
        }
        delete v;
        return M4G_OK;
    });
}

int m4g_vector_size(m4g_vector v, int64_t* n)
{
    return guard([&] {
        M4G_REQUIRE(v && n, M4G_ERR_INVALID, "m4g_vector_size: null argument");
        *n = v->n;
        return M4G_OK;
    });
}

void* m4g_vector_device_ptr(m4g_vector v) { return v ? v->ptr : nullptr; }

int m4g_vector_upload(m4g_vector v, const void* host)
{
    return guard([&] {
        M4G_REQUIRE(v && (host || v->n == 0), M4G_ERR_INVALID, "m4g_vector_upload: null argument");
        if (v->n)
            M4G_HIP(hipMemcpyAsync(v->ptr, host, v->bytes(), hipMemcpyHostToDevice, current_stream()));
        M4G_HIP(hipStreamSynchronize(current_stream()));
        return M4G_OK;
    });
}

int m4g_vector_download(m4g_vector v, void* host)
{
    return guard([&] {
        M4G_REQUIRE(v && (host || v->n == 0), M4G_ERR_INVALID, "m4g_vector_download: null argument");
        if (v->n)
            M4G_HIP(hipMemcpyAsync(host, v->ptr, v->bytes(), hipMemcpyDeviceToHost, current_stream()));
        M4G_HIP(hipStreamSynchronize(current_stream()));
        return M4G_OK;
    });
}

int m4g_vector_fill(m4g_vector v, double value)
{
    return guard([&] {
        M4G_REQUIRE(v, M4G_ERR_INVALID, "m4g_vector_fill: null vector");
        TraceRange tr(value == 0.0 ? kTraceSetZero : kTraceVecExpr);
        launch_fill(v->n, value, v->ptr, v->type, current_stream());
        return M4G_OK;
    });
}

int m4g_vector_copy(m4g_vector dst, m4g_vector src)
{
    TraceRange tr(kTraceVectorCopy);
    return guard([&] {
        M4G_REQUIRE(dst && src, M4G_ERR_INVALID, "m4g_vector_copy: null vector");
        M4G_REQUIRE(dst->n == src->n, M4G_ERR_SIZE, "m4g_vector_copy: incompatible size");
        if (dst->type == src->type) {
            if (dst->n)
                M4G_HIP(hipMemcpyAsync(dst->ptr, src->ptr, src->bytes(), hipMemcpyDeviceToDevice,
                                       current_stream()));
        } else {
            launch_convert(src->n, src->ptr, src->type, dst->ptr, dst->type, current_stream());
        }
        return M4G_OK;
    });
}

// ------------------------------- matrices ----------------------------------
int m4g_matrix_create(m4g_matrix* out, int64_t nrows, int64_t ncols, int64_t nnz, const void* starts,
                      const void* indices, const void* values, int index_type, int value_type)
{
    return guard([&] {
        M4G_REQUIRE(out && nrows >= 0 && ncols >= 0 && nnz >= 0 && starts && (nnz == 0 || (indices && values)),
                    M4G_ERR_INVALID, "m4g_matrix_create: bad argument");
        M4G_REQUIRE(value_type == M4G_F64 || value_type == M4G_F32, M4G_ERR_INVALID,
                    "m4g_matrix_create: unknown value type");
        M4G_REQUIRE(nnz < (int64_t(1) << 31) && nrows < (int64_t(1) << 31) && ncols < (int64_t(1) << 31),
                    M4G_ERR_INVALID, "m4g_matrix_create: dimensions exceed int32 device indices");
        auto A = std::make_unique<m4g_matrix_s>();
        A->nrows = nrows;
        A->ncols = ncols;
        A->nnz = nnz;
        A->input_value_type = value_type;
        widen_index(starts, nrows + 1, index_type, A->h_starts);
        widen_index(indices, nnz, index_type, A->h_indices);
        M4G_REQUIRE(A->h_starts[0] == 0 && A->h_starts[nrows] == nnz, M4G_ERR_INVALID,
                    "m4g_matrix_create: starts[0] must be 0 and starts[nrows] == nnz");
        for (int64_t r = 0; r < nrows; r++) {
            int64_t b = A->h_starts[r], e = A->h_starts[r + 1];
            M4G_REQUIRE(b <= e, M4G_ERR_INVALID, "m4g_matrix_create: starts not monotone");
            for (int64_t j = b; j < e; j++) {
                int64_t c = A->h_indices[j];
                M4G_REQUIRE(c >= 0 && c < ncols, M4G_ERR_INVALID, "m4g_matrix_create: column out of range");
                M4G_REQUIRE(j == b || A->h_indices[j - 1] < c, M4G_ERR_LOGIC,
                            "Matrix entries must be sorted for this.");
            }
        }
        std::vector<double> vals(nnz);
        if (value_type == M4G_F64) std::memcpy(vals.data(), values, nnz * sizeof(double));
        else
            for (int64_t j = 0; j < nnz; j++) vals[j] = static_cast<const float*>(values)[j];
        std::vector<int32_t> s32(A->h_starts.begin(), A->h_starts.end());
        std::vector<int32_t> i32(A->h_indices.begin(), A->h_indices.end());
        A->starts.reset(nrows + 1);
        A->indices.reset(nnz);
        A->values.reset(nnz);
        A->starts.upload(s32.data(), nrows + 1);
        A->indices.upload(i32.data(), nnz);
        A->values.upload(vals.data(), nnz);
        sell_build(*A);
        M4G_HIP(hipStreamSynchronize(current_stream()));
        *out = A.release();
        return M4G_OK;
    });
}

int m4g_matrix_destroy(m4g_matrix A)
{
    return guard([&] {
        delete A;
        return M4G_OK;
    });
}

int m4g_matrix_dims(m4g_matrix A, int64_t* nrows, int64_t* ncols, int64_t* nnz)
{
    return guard([&] {
        M4G_REQUIRE(A, M4G_ERR_INVALID, "m4g_matrix_dims: null matrix");
        if (nrows) *nrows = A->nrows;
        if (ncols) *ncols = A->ncols;
        if (nnz) *nnz = A->nnz;
        return M4G_OK;
    });
}

int m4g_matrix_download(m4g_matrix A, int64_t* starts, int64_t* indices, double* values)
{
    return guard([&] {
        M4G_REQUIRE(A, M4G_ERR_INVALID, "m4g_matrix_download: null matrix");
        if (starts) std::copy(A->h_starts.begin(), A->h_starts.end(), starts);
        if (indices) std::copy(A->h_indices.begin(), A->h_indices.end(), indices);
        if (values) A->values.download(values, A->nnz);
        return M4G_OK;
    });
}

int m4g_matrix_set_values(m4g_matrix A, const double* values, int values_device)
{
    return guard([&] {
        M4G_REQUIRE(A && (values || A->nnz == 0), M4G_ERR_INVALID, "m4g_matrix_set_values: null argument");
        if (A->nnz)
            M4G_HIP(hipMemcpyAsync(A->values.ptr, values, A->nnz * sizeof(double),
                                   values_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                   current_stream()));
        sell_refresh_values(*A);
        if (!values_device) M4G_HIP(hipStreamSynchronize(current_stream()));
        return M4G_OK;
    });
}

int m4g_matrix_transpose(m4g_matrix A, m4g_matrix* out)
{
    return guard([&] {
        M4G_REQUIRE(A && out, M4G_ERR_INVALID, "m4g_matrix_transpose: null argument");
        DevBuf<int64_t> pos;
        auto T = transpose_pattern(*A, pos);
        launch_gather_values(A->nnz, pos.ptr, A->values.ptr, T->values.ptr, current_stream());
        sell_build(*T);
        M4G_HIP(hipStreamSynchronize(current_stream()));
        *out = T.release();
        return M4G_OK;
    });
}

int m4g_pattern_gradient(m4g_matrix A, m4g_vector lambda, m4g_vector x, m4g_vector grad)
{
    return guard([&] {
        M4G_REQUIRE(A && lambda && x && grad, M4G_ERR_INVALID, "m4g_pattern_gradient: null argument");
        M4G_REQUIRE(lambda->type == M4G_F64 && x->type == M4G_F64 && grad->type == M4G_F64, M4G_ERR_INVALID,
                    "m4g_pattern_gradient: double vectors");
        M4G_REQUIRE(lambda->n == A->nrows && x->n == A->ncols && grad->n == A->nnz, M4G_ERR_SIZE,
                    "m4g_pattern_gradient: incompatible size");
        launch_pattern_gradient(*A, (const double*)lambda->ptr, (const double*)x->ptr, (double*)grad->ptr,
                                current_stream());
        return M4G_OK;
    });
}

// ------------------------------- level 1 -----------------------------------
static void check_f64(m4g_vector v, const char* who)
{
    M4G_REQUIRE(v, M4G_ERR_INVALID, std::string(who) + ": null vector");
    M4G_REQUIRE(v->type == M4G_F64, M4G_ERR_INVALID, std::string(who) + ": vector must be double");
}

int m4g_dot(m4g_vector x, m4g_vector y, double* result)
{
    return guard([&] {
        check_f64(x, "m4g_dot");
        check_f64(y, "m4g_dot");
        M4G_REQUIRE(x->n == y->n, M4G_ERR_SIZE, "m4g_dot: incompatible size");
        ReduceWs& ws = reduce_ws();
        double* out = ws.partials + kReduceBlocks * 16;
        hipStream_t s = current_stream();
        Mailbox mb(s);
        const unsigned q = mb.arm();
        launch_dot_post(x->n, (const double*)x->ptr, (const double*)y->ptr, out, mb.box(), q, s);
        *result = mb.wait(q)[0];
        return M4G_OK;
    });
}

int m4g_dot_noalloc(m4g_vector x, m4g_vector y, double* result)
{
    return guard([&] {
        check_f64(x, "m4g_dot_noalloc");
        check_f64(y, "m4g_dot_noalloc");
        M4G_REQUIRE(result, M4G_ERR_INVALID, "m4g_dot_noalloc: null result");
        M4G_REQUIRE(x->n == y->n, M4G_ERR_SIZE, "m4g_dot_noalloc: incompatible size");
        // set-up (m4g_set_device, or the first handle made on this device)
        // created the state and looked up the scratch; nothing is created here
        DeviceState* st = existing_device_state();
        M4G_REQUIRE(st && st->noalloc_scratch, M4G_ERR_INVALID,
                    "m4g_dot_noalloc: the device is not set up (call m4g_set_device first)");
        const int parts = reduce_partials(x->n);
        M4G_REQUIRE(parts <= kNoAllocResult, M4G_ERR_SIZE,
                    "m4g_dot_noalloc: more reduction partials than the scratch holds");
        hipStream_t s = tl_stream ? tl_stream : st->stream;
        double* out = st->noalloc_scratch + kNoAllocResult;
        // every partial the second stage reads is written by the first, and the
        // result slot by the second: nothing left from an earlier call is read
        launch_dot_with(x->n, (const double*)x->ptr, (const double*)y->ptr, st->noalloc_scratch, out, s);
        M4G_HIP(hipMemcpyAsync(st->pinned + 1, out, sizeof(double), hipMemcpyDeviceToHost, s));
        M4G_HIP(hipStreamSynchronize(s));
        *result = st->pinned[1];
        return M4G_OK;
    });
}

int m4g_unary_dot(m4g_vector x, double* result)
{
    return guard([&] {
        check_f64(x, "m4g_unary_dot");
        ReduceWs& ws = reduce_ws();
        double* out = ws.partials + kReduceBlocks * 16;
        hipStream_t s = current_stream();
        Mailbox mb(s);
        const unsigned q = mb.arm();
        launch_sumsq_post(x->n, (const double*)x->ptr, out, mb.box(), q, s);
        *result = mb.wait(q)[0];
        return M4G_OK;
    });
}

int m4g_two_norm(m4g_vector x, double* result)
{
    TraceRange tr(kTraceTwoNorm);
    int st = m4g_unary_dot(x, result);
    if (st == M4G_OK) *result = std::sqrt(*result);
    return st;
}

int m4g_axpy(double alpha, m4g_vector x, m4g_vector y)
{
    return guard([&] {
        check_f64(x, "m4g_axpy");
        check_f64(y, "m4g_axpy");
        M4G_REQUIRE(x->n == y->n, M4G_ERR_SIZE, "m4g_axpy: incompatible size");
        launch_axpy(x->n, alpha, (const double*)x->ptr, (double*)y->ptr, current_stream());
        return M4G_OK;
    });
}

int m4g_xpay(m4g_vector x, double beta, m4g_vector y)
{
    return guard([&] {
        check_f64(x, "m4g_xpay");
        check_f64(y, "m4g_xpay");
        M4G_REQUIRE(x->n == y->n, M4G_ERR_SIZE, "m4g_xpay: incompatible size");
        launch_xpay(x->n, (const double*)x->ptr, beta, (double*)y->ptr, current_stream());
        return M4G_OK;
    });
}

int m4g_axpy_unary_dot(double alpha, m4g_vector q, m4g_vector r, double* rr)
{
    return guard([&] {
        check_f64(q, "m4g_axpy_unary_dot");
        check_f64(r, "m4g_axpy_unary_dot");
        M4G_REQUIRE(q->n == r->n, M4G_ERR_SIZE, "m4g_axpy_unary_dot: incompatible size");
        ReduceWs& ws = reduce_ws();
        double* out = ws.partials + kReduceBlocks * 16;
        launch_axpy_sumsq(q->n, alpha, (const double*)q->ptr, (double*)r->ptr, out, current_stream());
        *rr = read_scalar(out);
        return M4G_OK;
    });
}

// ------------------------------- SpMV ---------------------------------------
static void check_spmv(m4g_matrix A, m4g_vector v, m4g_vector w, const char* who)
{
    M4G_REQUIRE(A, M4G_ERR_INVALID, std::string(who) + ": null matrix");
    check_f64(v, who);
    check_f64(w, who);
    M4G_REQUIRE(v->n == A->ncols && w->n == A->nrows, M4G_ERR_SIZE, std::string(who) + ": incompatible size");
    M4G_REQUIRE(v->ptr != w->ptr || A->nrows == 0, M4G_ERR_INVALID, std::string(who) + ": v and w alias");
}

int m4g_spmv(m4g_matrix A, m4g_vector v, m4g_vector w, int mode)
{
    return guard([&] {
        check_spmv(A, v, w, "m4g_spmv");
        M4G_REQUIRE(mode == M4G_ASSIGN || mode == M4G_PLUS || mode == M4G_MINUS, M4G_ERR_INVALID,
                    "m4g_spmv: unknown mode");
        launch_spmv(*A, (const double*)v->ptr, (double*)w->ptr, mode, nullptr, current_stream());
        return M4G_OK;
    });
}

int m4g_residual(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_vector r)
{
    return guard([&] {
        check_spmv(A, x, r, "m4g_residual");
        check_f64(b, "m4g_residual");
        M4G_REQUIRE(b->n == A->nrows, M4G_ERR_SIZE, "m4g_residual: incompatible size");
        launch_spmv(*A, (const double*)x->ptr, (double*)r->ptr, M4G_ASSIGN, (const double*)b->ptr,
                    current_stream());
        return M4G_OK;
    });
}

int m4g_spmv_dot(m4g_matrix A, m4g_vector p, m4g_vector q, double* pq)
{
    return guard([&] {
        check_spmv(A, p, q, "m4g_spmv_dot");
        M4G_REQUIRE(A->nrows == A->ncols, M4G_ERR_NOT_SQUARE, "m4g_spmv_dot: matrix not square");
        ReduceWs& ws = reduce_ws();
        double* out = ws.partials + kReduceBlocks * 16;
        launch_spmv_dot(*A, (const double*)p->ptr, (double*)q->ptr, out, current_stream());
        *pq = read_scalar(out);
        return M4G_OK;
    });
}

}  // extern "C"
