// Internal declarations of the gfx950 sparse-solve library (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mtl4gpu.h"
#include "trace.hpp"

namespace m4g {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

void set_last_error(const std::string& msg);

#define M4G_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::m4g::Error(M4G_ERR_HIP, std::string(#call " failed: ") +            \
                                                hipGetErrorString(e_) + " at " +        \
                                                __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

#define M4G_REQUIRE(cond, status, msg)                                                  \
    do {                                                                                \
        if (!(cond)) throw ::m4g::Error((status), (msg));                               \
    } while (0)

// Runs f, converting exceptions into status codes (the C ABI never throws).
template <typename F>
int guard(F&& f)
{
    try {
        return f();
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return M4G_ERR_HIP;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return M4G_ERR_INVALID;
    }
}

// ---------------------------------------------------------------------------
// device and stream state
// ---------------------------------------------------------------------------
struct HostPost;
struct DeviceState {
    int device = 0;
    hipStream_t stream = nullptr;      // library stream (default priority)
    int num_cus = 0;
    std::mutex pool_mutex;
    std::multimap<size_t, void*> pool; // freed buffers by size, reused exactly
    double* pinned = nullptr;          // small pinned host staging buffer (scalars)
    double* noalloc_scratch = nullptr; // this device's copy of m4g_dot_noalloc's scratch array
    std::mutex mail_mutex;
    std::vector<std::pair<HostPost*, unsigned>> mailboxes;   // idle mailboxes, next sequence number
};

DeviceState& device_state();          // state of the calling thread's device (created on first use)
DeviceState* existing_device_state(); // the same, or nullptr when it has not been set up
hipStream_t current_stream();          // stream for this thread's work

void* pool_alloc(size_t bytes);
void pool_free(void* p, size_t bytes);

// RAII device buffer from the pool.
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { reset(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) { release(); ptr = o.ptr; count = o.count; o.ptr = nullptr; o.count = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void reset(size_t n)
    {
        release();
        count = n;
        ptr = n ? static_cast<T*>(pool_alloc(n * sizeof(T))) : nullptr;
    }
    void release()
    {
        if (ptr) pool_free(ptr, count * sizeof(T));
        ptr = nullptr;
        count = 0;
    }
    void upload(const T* host, size_t n)
    {
        if (n) M4G_HIP(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, current_stream()));
    }
    void download(T* host, size_t n) const
    {
        if (n) {
            M4G_HIP(hipMemcpyAsync(host, ptr, n * sizeof(T), hipMemcpyDeviceToHost, current_stream()));
            M4G_HIP(hipStreamSynchronize(current_stream()));
        }
    }
};

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct Sell;   // sliced-ELL copy used by SpMV (spmv.hip)

}  // namespace m4g

struct m4g_vector_s {
    int64_t n = 0;
    int type = M4G_F64;
    void* ptr = nullptr;
    bool owned = false;
    size_t bytes() const { return size_t(n) * (type == M4G_F64 ? 8 : 4); }
};

namespace m4g {

// Sliced ELL: rows in slices of 64 (one wavefront); within a slice, entry k of
// all 64 rows is contiguous, so lane l reads entry k of row 64*s + l with one
// coalesced load.  Column storage per slice: int16 column − row ("narrow") or
// int32 columns ("wide"), see spmv_dev.hpp.
struct Sell {
    int64_t n = 0, nslices = 0, padded = 0;
    DevBuf<int64_t> slice_ptr;   // [nslices+1] start of each slice (entries)
    DevBuf<int64_t> slice_idx;   // [nslices] byte offset of the slice's indices << 1 | wide
    DevBuf<uint8_t> idx;         // index bytes of all slices
    int64_t narrow_slices = 0;
    DevBuf<double> vals;         // [padded]
    DevBuf<float> vals32;        // [padded] (lazily built for FP32 paths)
    DevBuf<int64_t> csr_pos;     // [padded] CSR index of each entry (-1 for padding)
    int max_width = 0;
};

}  // namespace m4g

struct m4g_matrix_s {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    // host copy of the pattern (analysis, schedules, downloads)
    std::vector<int64_t> h_starts, h_indices;
    // device CSR (int32 indices), values in double
    m4g::DevBuf<int32_t> starts, indices;
    m4g::DevBuf<double> values;
    m4g::Sell sell;
    int input_value_type = M4G_F64;
    // reduction workspace of the fused SpMV + dot (one partial per workgroup)
    m4g::DevBuf<double> red_partials;
    m4g::DevBuf<unsigned> red_counter;
    // trans(A) for the adjoint products of bicg / qmr: pattern, SELL layout and
    // the map to A's entries built on first use, values gathered at each use
    std::unique_ptr<m4g_matrix_s> tr;
    m4g::DevBuf<int64_t> tr_pos;
};

struct m4g_pc_s;   // defined in pc.hip

namespace m4g {

// ---------------------------------------------------------------------------
// kernel launchers shared between translation units
// ---------------------------------------------------------------------------
void sell_build(m4g_matrix_s& A);                   // spmv.hip
// trans(A) with A's current values (runtime.hip): the pattern is built on the
// first call and kept with A, the values are gathered from A on stream s
const m4g_matrix_s& transpose_of(m4g_matrix_s& A, hipStream_t s);
void sell_refresh_values(m4g_matrix_s& A);          // spmv.hip
void sell_build_f32(m4g_matrix_s& A);               // spmv.hip

// y = A*x with mode (M4G_ASSIGN/PLUS/MINUS), or r = b - A*x when b != nullptr.
void launch_spmv(const m4g_matrix_s& A, const double* x, double* y, int mode, const double* b,
                 hipStream_t s);
// q = A*p and per-block partial sums of p.q into partials[]; the last block
// writes the total to *out.  Returns nothing; caller owns partials/counter.
void launch_spmv_dot(const m4g_matrix_s& A, const double* p, double* q, double* out, hipStream_t s);

// Reductions into a device scalar (deterministic two-stage).
void launch_dot(int64_t n, const double* x, const double* y, double* out, hipStream_t s);
// dot with caller-provided partials (room for reduce_partials(n) doubles)
int reduce_partials(int64_t n);
void launch_dot_with(int64_t n, const double* x, const double* y, double* partials, double* out, hipStream_t s);
void launch_sumsq(int64_t n, const double* x, double* out, hipStream_t s);
void launch_axpy(int64_t n, double alpha, const double* x, double* y, hipStream_t s);
void launch_xpay(int64_t n, const double* x, double beta, double* y, hipStream_t s);
void launch_axpy_sumsq(int64_t n, double alpha, const double* q, double* r, double* out, hipStream_t s);
void launch_fill(int64_t n, double value, void* x, int type, hipStream_t s);
void launch_sum_partials(const double* partials, int nblocks, double* out, hipStream_t s);
// pipelined SpMV: its MAXW for A (0: not applicable) and its grid
int sell_pipe_width(const m4g_matrix_s& A);
int sell_pipe_grid();
void launch_pattern_gradient(const m4g_matrix_s& A, const double* lambda, const double* x, double* grad,
                             hipStream_t s);
void launch_gather_values(int64_t count, const int64_t* pos, const double* src, double* dst, hipStream_t s);
void launch_convert(int64_t n, const void* src, int src_type, void* dst, int dst_type, hipStream_t s);

// Reduction workspace: per-device partials and counters.
struct ReduceWs {
    double* partials = nullptr;      // [kMaxBlocks * 4]
    unsigned* counters = nullptr;    // [kCounters]
};
ReduceWs& reduce_ws();
constexpr int kReduceBlocks = 1024;  // blocks of the first reduction stage
constexpr int kCounters = 64;

// Host read of a device double once the stream's work before it is done
// (posted through a mailbox, below).
double read_scalar(const double* d);

// Scalars that a kernel posts straight into pinned, host-coherent memory,
// followed by a sequence number.  The host spins on the number instead of
// copying the value back and synchronising the stream: about 6.5 us from a
// small kernel's launch to its value on the host, against 14 us for a copy
// and a stream synchronisation.
constexpr int kPostMax = 64;
struct HostPost {
    double val[kPostMax];
    unsigned seq;
};

// One of the device's mailboxes, leased for the lifetime of a solve on one
// stream.  arm() gives the number the next post carries; wait() returns the
// posted values once that number has arrived (after 50 ms of spinning it
// synchronises the stream instead, which also reports a failed kernel).  A
// mailbox goes back to the pool only once every armed post has arrived.
class Mailbox {
public:
    explicit Mailbox(hipStream_t s);
    ~Mailbox();
    Mailbox(const Mailbox&) = delete;
    Mailbox& operator=(const Mailbox&) = delete;
    HostPost* box() const { return box_; }
    unsigned arm() { return ++armed_; }
    const double* wait(unsigned seq);

private:
    HostPost* box_ = nullptr;
    unsigned armed_ = 0, seen_ = 0;
    hipStream_t s_;
};

// dot / squared norm into *out, the result also posted into box[0] with seq
void launch_dot_post(int64_t n, const double* x, const double* y, double* out, HostPost* box, unsigned seq,
                     hipStream_t s);
void launch_sumsq_post(int64_t n, const double* x, double* out, HostPost* box, unsigned seq, hipStream_t s);
// posts src[0..count) (count <= kPostMax, device memory) into box with seq
void launch_post(const double* src, int count, HostPost* box, unsigned seq, hipStream_t s);

}  // namespace m4g
