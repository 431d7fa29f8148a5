// C++ compatibility layer: the reference's call surface for the sparse solve
// path (mtl::compressed2D, mtl::dense_vector, itl::cg, itl::bicgstab,
// itl::basic_iteration / cyclic_iteration / noisy_iteration,
// itl::pc::{identity, diagonal, ic_0, ilu_0}) over the C ABI of libmtl4gpu.
//
// Host code written against MTL4 for this path compiles unchanged when the
// include path points at include/mtl4gpu (which provides
// <boost/numeric/mtl/mtl.hpp> and <boost/numeric/itl/itl.hpp>) and the program
// links libmtl4gpu.so.  Matrices and vectors live on the GPU; element access
// through operator[] works on a host mirror that is synchronised on demand.
//
// Interfaces mirrored (paths relative to the reference checkout):
//   itl/krylov/cg.hpp:39-179, itl/krylov/bicgstab.hpp:27-112,
//   itl/krylov/base_solver.hpp:24-84, itl/iteration/*.hpp,
//   itl/pc/{ic_0,ilu_0,ilu,diagonal,identity,solver,is_identity}.hpp,
//   mtl/matrix/compressed2D.hpp:258-653 (pointer constructor :355-364),
//   mtl/matrix/laplacian_setup.hpp:25-44, mtl/vector/dense_vector.hpp,
//   mtl/operation/{dot,two_norm,unary_dot}.hpp, mtl/utility/exception.hpp.
#ifndef MTL4GPU_HPP
#define MTL4GPU_HPP

#include <algorithm>
#include <thread>
#include <atomic>
#include <type_traits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <initializer_list>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../mtl4gpu.h"
#include "../mtl4gpu_gen.h"
#include "../mtl4gpu_io.h"
#include "../mtl4gpu_formats.h"

namespace mtl {

// ---------------------------------------------------------------------------
// exceptions (mtl/utility/exception.hpp:95-230)
// ---------------------------------------------------------------------------
struct runtime_error : std::runtime_error {
    explicit runtime_error(const char* s = "Run-time error") : std::runtime_error(s) {}
};
struct logic_error : std::logic_error {
    explicit logic_error(const char* s = "Logic error") : std::logic_error(s) {}
};
struct index_out_of_range : logic_error {
    explicit index_out_of_range(const char* s = "Index out of range") : logic_error(s) {}
};
struct incompatible_size : logic_error {
    explicit incompatible_size(const char* s = "Arguments have incompatible size.") : logic_error(s) {}
};
struct matrix_not_square : logic_error {
    explicit matrix_not_square(const char* s = "Matrix must be square for this operation.") : logic_error(s) {}
};
struct missing_diagonal : logic_error {
    explicit missing_diagonal(const char* s = "Diagonal entry missing or not where it belongs to.") : logic_error(s) {}
};
struct domain_error : std::domain_error {
    explicit domain_error(const char* s = "MTL4 domain error.") : std::domain_error(s) {}
};
struct matrix_singular : domain_error {   // exception.hpp:172-177
    explicit matrix_singular(const char* s = "Matrix singular in solver.") : domain_error(s) {}
};
struct io_error : runtime_error {
    explicit io_error(const char* s = "I/O error") : runtime_error(s) {}
};
struct file_not_found : io_error {   // exception.hpp:232-236
    explicit file_not_found(const char* s = "File not found") : io_error(s) {}
};

#define MTL_THROW_IF(Test, Exception) \
    do {                              \
        if (Test) throw Exception;    \
    } while (0)
#define MTL_THROW(Exception) throw Exception

namespace detail {
// Maps a C-ABI status to the reference's exception types.
void throw_status(int st, const char* where);
}  // namespace detail

}  // namespace mtl

namespace itl {
// itl/utility/exception.hpp:22-31
struct unexpected_orthogonality : mtl::runtime_error {
    unexpected_orthogonality() : mtl::runtime_error("Unexpected orthogonality, e.g. in BiCGStab") {}
};
}  // namespace itl

namespace mtl {
namespace detail {
inline void throw_status(int st, const char* where)
{
    if (st == M4G_OK || st == M4G_MAX_ITER) return;
    std::string msg = std::string(where) + ": " + m4g_last_error();
    switch (st) {
    case M4G_ERR_NOT_SQUARE: throw matrix_not_square();
    case M4G_ERR_MISSING_DIAGONAL: throw missing_diagonal();
    case M4G_ERR_ORTHOGONALITY: throw itl::unexpected_orthogonality();
    case M4G_ERR_LOGIC: throw logic_error(m4g_last_error());
    case M4G_ERR_SIZE: throw incompatible_size();
    case M4G_ERR_FILE: throw file_not_found(m4g_last_error());
    case M4G_ERR_FORMAT: throw runtime_error(m4g_last_error());
    case M4G_ERR_SINGULAR: throw matrix_singular();
    default: throw runtime_error(msg.c_str());
    }
}

template <typename Value> struct value_code;
template <> struct value_code<double> { static const int code = M4G_F64; };
template <> struct value_code<float> { static const int code = M4G_F32; };
}  // namespace detail

namespace mat { template <typename Value> class compressed2D; }
namespace vec { template <typename Value> class dense_vector; }

namespace io {
// A Matrix Market file name, read on construction or assignment of a matrix:
// `compressed2D<double> A(io::matrix_market("a.mtx"))`, `A = io::matrix_market(f)`
// (matrix_file.hpp, matrix_market.hpp)
class matrix_market {
  public:
    explicit matrix_market(const std::string& file_name) : file_name_(file_name) {}
    const std::string& file_name() const { return file_name_; }
    template <typename Matrix> void read_into(Matrix& A) const;

  private:
    std::string file_name_;
};
}  // namespace io

// ---------------------------------------------------------------------------
// lazy expressions of the path
// ---------------------------------------------------------------------------
template <typename Matrix, typename Vector>
struct mat_cvec_times_expr {                 // A * v (mat_cvec_times_expr.hpp:23-36)
    const Matrix& A;
    const Vector& v;
};
template <typename Vector>
struct scaled_vector {                       // alpha * v
    double alpha;
    const Vector& v;
};
template <typename Vector, typename Matrix>
struct residual_expr {                       // b - A*x
    const Vector& b;
    mat_cvec_times_expr<Matrix, Vector> Ax;
};
template <typename Vector, typename Matrix>
struct product_minus_expr {                  // A*x - b
    mat_cvec_times_expr<Matrix, Vector> Ax;
    const Vector& b;
};
template <typename Vector>
struct vec_sum_expr {                        // x + alpha*y (alpha = -1 for x - y)
    const Vector& x;
    double alpha;
    const Vector& y;
};

namespace detail {
// true for the library's matrix types, which carry a device handle
template <typename M, typename = void>
struct has_handle : std::false_type {};
template <typename M>
struct has_handle<M, std::void_t<decltype(std::declval<const M&>().handle())>> : std::true_type {};
}  // namespace detail

// assign modes of the operator protocol (mtl/operation/assign_mode.hpp:22-135)
namespace assign {
struct assign_sum {
    template <typename T, typename U> static T& apply(T& x, const U& y) { return x = y; }
};
struct plus_sum {
    template <typename T, typename U> static T& apply(T& x, const U& y) { return x += y; }
};
struct minus_sum {
    template <typename T, typename U> static T& apply(T& x, const U& y) { return x -= y; }
};
}  // namespace assign

namespace detail {
template <typename Assign> struct assign_code;
template <> struct assign_code<assign::assign_sum> { static const int value = M4G_ASSIGN; };
template <> struct assign_code<assign::plus_sum> { static const int value = M4G_PLUS; };
template <> struct assign_code<assign::minus_sum> { static const int value = M4G_MINUS; };
}  // namespace detail

// algebraic shapes: a user operator declares ashape_aux<Op>::type = nonscal
// (poisson2D_dirichlet.hpp:96-101); the drop-in only needs the names
namespace ashape {
struct scal {};
struct nonscal {};
template <typename T> struct ashape_aux { typedef nonscal type; };
}  // namespace ashape

namespace vec {

// A * v for a matrix-free operator: the product is computed where it goes,
// through A.mult(v, w, Assign) (mtl/vector/mat_cvec_multiplier.hpp:25-69)
template <typename Matrix, typename VectorIn>
struct mat_cvec_multiplier {
    mat_cvec_multiplier(const Matrix& A, const VectorIn& v) : A(A), v(v) {}
    template <typename VectorOut> void assign_to(VectorOut& w) const { A.mult(v, w, mtl::assign::assign_sum()); }
    template <typename VectorOut> void increment_it(VectorOut& w) const { A.mult(v, w, mtl::assign::plus_sum()); }
    template <typename VectorOut> void decrement_it(VectorOut& w) const { A.mult(v, w, mtl::assign::minus_sum()); }
    const Matrix& A;
    const VectorIn& v;
};

// ---------------------------------------------------------------------------
// dense_vector: device-resident, with an on-demand host mirror
// ---------------------------------------------------------------------------
template <typename Value = double>
class dense_vector {
  public:
    typedef Value value_type;
    typedef std::size_t size_type;
    typedef dense_vector self;

    dense_vector() = default;
    explicit dense_vector(size_type n) { change_resource(n); }
    dense_vector(size_type n, Value v) { change_resource(n); *this = v; }
    template <std::size_t N>
    explicit dense_vector(const Value (&a)[N])
    {
        change_resource(N);
        host_ = std::vector<Value>(a, a + N);
        state_ = host_valid;
    }
    dense_vector(const dense_vector& o) { copy_from(o); }
    dense_vector(dense_vector&& o) noexcept { swap(o); }
    template <typename Matrix>
    dense_vector(const mat_cvec_times_expr<Matrix, dense_vector>& e) { *this = e; }
    template <typename Matrix>
    dense_vector(const residual_expr<dense_vector, Matrix>& e) { *this = e; }
    template <typename Matrix>
    dense_vector(const product_minus_expr<dense_vector, Matrix>& e) { *this = e; }
    dense_vector(const vec_sum_expr<dense_vector>& e) { *this = e; }
    ~dense_vector()
    {
        if (h_) m4g_vector_destroy(h_);
    }

    dense_vector& operator=(const dense_vector& o)
    {
        if (this != &o) copy_from(o);
        return *this;
    }
    dense_vector& operator=(dense_vector&& o) noexcept
    {
        swap(o);
        return *this;
    }
    // comma initialiser: v = 1.0, 2.0, 3.0;  (v = s alone sets every element to s)
    struct comma_init {
        dense_vector& v;
        size_type pos;
        comma_init& operator,(Value x)
        {
            v[pos++] = x;
            return *this;
        }
    };
    comma_init operator=(Value v)
    {
        check(m4g_vector_fill(dev(), double(v)), "dense_vector fill");
        state_ = device_valid;
        return comma_init{*this, 1};
    }
    // w = A * v (crtp_base_vector.hpp:93-104)
    template <typename Matrix>
    dense_vector& operator=(const mat_cvec_times_expr<Matrix, dense_vector>& e)
    {
        checked_change_resource(num_rows(e.A));
        check(e.A.mult_device(e.v.dev(), dev(), M4G_ASSIGN), "A * v");
        state_ = device_valid;
        return *this;
    }
    template <typename Matrix>
    dense_vector& operator+=(const mat_cvec_times_expr<Matrix, dense_vector>& e)
    {
        check(e.A.mult_device(e.v.dev(), dev(), M4G_PLUS), "+= A * v");
        state_ = device_valid;
        return *this;
    }
    template <typename Matrix>
    dense_vector& operator-=(const mat_cvec_times_expr<Matrix, dense_vector>& e)
    {
        check(e.A.mult_device(e.v.dev(), dev(), M4G_MINUS), "-= A * v");
        state_ = device_valid;
        return *this;
    }
    // r = b - A*x (compute_summand.hpp:42-60)
    template <typename Matrix>
    dense_vector& operator=(const residual_expr<dense_vector, Matrix>& e)
    {
        checked_change_resource(num_rows(e.Ax.A));
        check(m4g_residual(e.Ax.A.handle(), e.Ax.v.dev(), e.b.dev(), dev()), "b - A * x");
        state_ = device_valid;
        return *this;
    }
    // r = A*x - b: w = A*x, then w -= b
    template <typename Matrix>
    dense_vector& operator=(const product_minus_expr<dense_vector, Matrix>& e)
    {
        dense_vector tmp(num_rows(e.Ax.A));
        check(m4g_spmv(e.Ax.A.handle(), e.Ax.v.dev(), tmp.dev(), M4G_ASSIGN), "A * x - b");
        check(m4g_axpy(-1.0, e.b.dev(), tmp.dev()), "A * x - b");
        tmp.state_ = device_valid;
        swap(tmp);
        return *this;
    }
    // z = x + alpha*y
    dense_vector& operator=(const vec_sum_expr<dense_vector>& e)
    {
        dense_vector tmp(e.x);
        check(m4g_axpy(e.alpha, e.y.dev(), tmp.dev()), "x + alpha*y");
        tmp.state_ = device_valid;
        swap(tmp);
        return *this;
    }
    dense_vector& operator+=(const scaled_vector<dense_vector>& e)
    {
        check(m4g_axpy(e.alpha, e.v.dev(), dev()), "+= alpha*v");
        state_ = device_valid;
        return *this;
    }
    dense_vector& operator-=(const scaled_vector<dense_vector>& e)
    {
        check(m4g_axpy(-e.alpha, e.v.dev(), dev()), "-= alpha*v");
        state_ = device_valid;
        return *this;
    }
    dense_vector& operator+=(const dense_vector& y) { return *this += scaled_vector<dense_vector>{1.0, y}; }
    dense_vector& operator-=(const dense_vector& y) { return *this -= scaled_vector<dense_vector>{1.0, y}; }
    dense_vector& operator*=(Value s)
    {
        dense_vector z(size(), Value(0));
        check(m4g_xpay(z.dev(), double(s), dev()), "*= s");   // y = 0 + s*y
        state_ = device_valid;
        return *this;
    }
    // assigner protocol (pc::solver): y = solve(P, x) calls P.solve(x, y) (pc/solver.hpp:33-48)
    template <typename Assigner, typename = decltype(std::declval<const Assigner&>().assign_to(std::declval<dense_vector&>()))>
    dense_vector& operator=(const Assigner& a)
    {
        a.assign_to(*this);
        return *this;
    }
    template <typename Assigner, typename = decltype(std::declval<const Assigner&>().assign_to(std::declval<dense_vector&>())),
              typename = void>
    dense_vector(const Assigner& a)
    {
        a.assign_to(*this);
    }
    // incrementer / decrementer protocol (w += A * v, w -= A * v for operators)
    template <typename Incr, typename = decltype(std::declval<const Incr&>().increment_it(std::declval<dense_vector&>()))>
    dense_vector& operator+=(const Incr& a)
    {
        a.increment_it(*this);
        return *this;
    }
    template <typename Decr, typename = decltype(std::declval<const Decr&>().decrement_it(std::declval<dense_vector&>()))>
    dense_vector& operator-=(const Decr& a)
    {
        a.decrement_it(*this);
        return *this;
    }

    size_type size() const { return size_type(n_); }
    void change_resource(size_type n)
    {
        if (h_ && size_type(n_) == n) return;
        if (h_) m4g_vector_destroy(h_);
        h_ = nullptr;
        n_ = int64_t(n);
        check(m4g_vector_create(&h_, n_, detail::value_code<Value>::code), "dense_vector");
        host_.assign(n, Value(0));
        state_ = host_valid;   // fresh storage: the mirror is authoritative
    }
    void checked_change_resource(size_type n)
    {
        if (!h_ || size() == 0) change_resource(n);
        else MTL_THROW_IF(size() != n, incompatible_size());
    }
    void checked_change_resource(const dense_vector& o) { checked_change_resource(o.size()); }

    // element access on the host mirror
    Value& operator[](size_type i)
    {
        to_host();
        state_ = host_valid;
        return host_[i];
    }
    const Value& operator[](size_type i) const
    {
        const_cast<dense_vector*>(this)->to_host();
        return host_[i];
    }
    Value& operator()(size_type i) { return (*this)[i]; }
    const Value& operator()(size_type i) const { return (*this)[i]; }

    // device handle for the C ABI (uploads a pending host mirror)
    m4g_vector dev() const
    {
        dense_vector* self_ = const_cast<dense_vector*>(this);
        if (!h_) self_->change_resource(0);
        if (state_ == host_valid) {
            check(m4g_vector_upload(h_, host_.data()), "dense_vector upload");
            self_->state_ = both_valid;
        }
        return h_;
    }
    void device_modified() { state_ = device_valid; }
    void swap(dense_vector& o) noexcept
    {
        std::swap(h_, o.h_);
        std::swap(n_, o.n_);
        std::swap(host_, o.host_);
        std::swap(state_, o.state_);
    }

  private:
    enum State { host_valid, device_valid, both_valid };
    void to_host()
    {
        if (state_ == device_valid) {
            host_.resize(n_);
            check(m4g_vector_download(h_, host_.data()), "dense_vector download");
            state_ = both_valid;
        }
    }
    void copy_from(const dense_vector& o)
    {
        change_resource(o.size());
        check(m4g_vector_copy(h_, o.dev()), "dense_vector copy");
        state_ = device_valid;
    }
    static void check(int st, const char* where) { detail::throw_status(st, where); }

    m4g_vector h_ = nullptr;
    int64_t n_ = 0;
    std::vector<Value> host_;
    State state_ = host_valid;
};

template <typename V> inline std::size_t size(const dense_vector<V>& v) { return v.size(); }
template <typename V> inline std::size_t num_rows(const dense_vector<V>& v) { return v.size(); }

template <typename V>
inline std::ostream& operator<<(std::ostream& os, const dense_vector<V>& v)
{
    os << "[";
    for (std::size_t i = 0; i < v.size(); i++) os << v[i] << (i + 1 < v.size() ? "," : "");
    return os << "]";
}

// one_norm(v) = sum |v_i| (mtl/operation/one_norm.hpp), summed on the host mirror
template <typename V>
inline V one_norm(const dense_vector<V>& v)
{
    V s = 0;
    for (std::size_t i = 0; i < v.size(); i++) s += std::abs(v[i]);
    return s;
}

// iota(v, offset): v[i] = i + offset
template <typename V>
inline void iota(dense_vector<V>& v, V offset = V(0))
{
    for (std::size_t i = 0; i < v.size(); i++) v[i] = V(i) + offset;
}

// dot(x, y) (dot.hpp:226-232), two_norm (two_norm.hpp:52-57), unary_dot (unary_dot.hpp:56-59)
template <typename V>
inline V dot(const dense_vector<V>& x, const dense_vector<V>& y)
{
    double r = 0;
    detail::throw_status(m4g_dot(x.dev(), y.dev(), &r), "dot");
    return V(r);
}
template <typename V>
inline V two_norm(const dense_vector<V>& x)
{
    double r = 0;
    detail::throw_status(m4g_two_norm(x.dev(), &r), "two_norm");
    return V(r);
}
template <typename V>
inline V unary_dot(const dense_vector<V>& x)
{
    double r = 0;
    detail::throw_status(m4g_unary_dot(x.dev(), &r), "unary_dot");
    return V(r);
}

template <typename V>
inline scaled_vector<dense_vector<V>> operator*(double a, const dense_vector<V>& v) { return {a, v}; }
template <typename V>
inline vec_sum_expr<dense_vector<V>> operator-(const dense_vector<V>& x, const dense_vector<V>& y) { return {x, -1.0, y}; }
template <typename V>
inline vec_sum_expr<dense_vector<V>> operator+(const dense_vector<V>& x, const dense_vector<V>& y) { return {x, 1.0, y}; }
template <typename V>
inline vec_sum_expr<dense_vector<V>> operator+(const dense_vector<V>& x, const scaled_vector<dense_vector<V>>& y)
{
    return {x, y.alpha, y.v};
}
template <typename V>
inline vec_sum_expr<dense_vector<V>> operator-(const dense_vector<V>& x, const scaled_vector<dense_vector<V>>& y)
{
    return {x, -y.alpha, y.v};
}

}  // namespace vec

using vec::dense_vector;
using vec::dot;
using vec::iota;
using vec::two_norm;
using vec::unary_dot;

namespace mat {

// ---------------------------------------------------------------------------
// compressed2D: device-resident CSR with a host copy of the arrays
// ---------------------------------------------------------------------------
template <typename Value = double>
class compressed2D {
  public:
    typedef Value value_type;
    typedef std::size_t size_type;

    compressed2D() = default;
    compressed2D(size_type m, size_type n) : nrows_(m), ncols_(n), starts_(m + 1, 0) {}
    // compressed2D(m, n, nnz, starts, indices, data): copies (compressed2D.hpp:355-364)
    template <typename Index>
    compressed2D(size_type m, size_type n, size_type nnz, const Index* starts, const Index* indices, const Value* data)
        : nrows_(m), ncols_(n), starts_(starts, starts + m + 1), indices_(indices, indices + nnz), data_(data, data + nnz)
    {
        upload();
    }
    compressed2D(const compressed2D& o)
        : nrows_(o.nrows_), ncols_(o.ncols_), starts_(o.starts_), indices_(o.indices_), data_(o.data_)
    {
        if (o.h_) upload();
    }
    compressed2D(const io::matrix_market& f) { f.read_into(*this); }
    compressed2D& operator=(const io::matrix_market& f)
    {
        f.read_into(*this);
        return *this;
    }
    compressed2D& operator=(const compressed2D& o)
    {
        if (this != &o) {
            nrows_ = o.nrows_;
            ncols_ = o.ncols_;
            starts_ = o.starts_;
            indices_ = o.indices_;
            data_ = o.data_;
            upload();
        }
        return *this;
    }
    ~compressed2D()
    {
        if (h_) m4g_matrix_destroy(h_);
    }

    void change_dim(size_type m, size_type n)
    {
        nrows_ = m;
        ncols_ = n;
        starts_.assign(m + 1, 0);
        indices_.clear();
        data_.clear();
        release();
    }
    // A = s sets the diagonal to s (T-mtl/fuse_test.cpp:33-40)
    compressed2D& operator=(Value s)
    {
        starts_.assign(nrows_ + 1, 0);
        indices_.clear();
        data_.clear();
        if (s != Value(0))
            for (size_type r = 0; r < std::min(nrows_, ncols_); r++) {
                indices_.push_back(int64_t(r));
                data_.push_back(s);
                starts_[r + 1] = int64_t(indices_.size());
            }
        for (size_type r = std::min(nrows_, ncols_); r < nrows_; r++) starts_[r + 1] = int64_t(indices_.size());
        upload();
        return *this;
    }

    // host-side construction from sorted rows (used by the inserter and generators)
    void assign_csr(std::vector<int64_t> S, std::vector<int64_t> I, std::vector<Value> D)
    {
        starts_ = std::move(S);
        indices_ = std::move(I);
        data_ = std::move(D);
        upload();
    }

    size_type nnz() const { return size_type(indices_.size()); }
    size_type num_rows() const { return nrows_; }
    size_type num_cols() const { return ncols_; }
    // raw views (compressed2D.hpp:609-625); indices are 64-bit like the
    // default size_type.  A non-const view may be written through, so it
    // drops the device copy, which is rebuilt from the host rows on next use.
    const std::vector<int64_t>& ref_major() const { return starts_; }
    const std::vector<int64_t>& ref_minor() const { return indices_; }
    std::vector<int64_t>& ref_major()
    {
        release();
        return starts_;
    }
    std::vector<int64_t>& ref_minor()
    {
        release();
        return indices_;
    }
    const int64_t* address_major() const { return starts_.data(); }
    const int64_t* address_minor() const { return indices_.data(); }
    const Value* address_data() const { return data_.data(); }
    int64_t* address_major()
    {
        release();
        return starts_.data();
    }
    int64_t* address_minor()
    {
        release();
        return indices_.data();
    }
    Value* address_data()
    {
        release();
        return data_.data();
    }
    const std::vector<Value>& data() const { return data_; }

    // A[r][c]: 0 when absent (compressed2D.hpp:536-547)
    struct row_proxy {
        const compressed2D& A;
        size_type r;
        Value operator[](size_type c) const { return A(r, c); }
    };
    row_proxy operator[](size_type r) const { return {*this, r}; }
    Value operator()(size_type r, size_type c) const
    {
        auto b = indices_.begin() + starts_[r], e = indices_.begin() + starts_[r + 1];
        auto p = std::lower_bound(b, e, int64_t(c));
        return (p != e && *p == int64_t(c)) ? data_[p - indices_.begin()] : Value(0);
    }

    m4g_matrix handle() const
    {
        if (!h_) const_cast<compressed2D*>(this)->upload();
        return h_;
    }
    // w (=, +=, -=) A * v on the device
    int mult_device(m4g_vector v, m4g_vector w, int mode) const { return m4g_spmv(handle(), v, w, mode); }

    // the host rows, as the inserter and the format conversions read them
    const std::vector<int64_t>& host_starts() const { return starts_; }
    const std::vector<int64_t>& host_indices() const { return indices_; }
    const std::vector<Value>& host_values() const { return data_; }

    // copy from another matrix type of this library (ell_matrix, sparse_banded)
    template <typename Src, typename = decltype(std::declval<const Src&>().host_starts()),
              typename = typename std::enable_if<!std::is_same<Src, compressed2D>::value>::type>
    explicit compressed2D(const Src& src) : nrows_(src.num_rows()), ncols_(src.num_cols())
    {
        assign_csr(src.host_starts(), src.host_indices(),
                   std::vector<Value>(src.host_values().begin(), src.host_values().end()));
    }
    template <typename Src, typename = decltype(std::declval<const Src&>().host_starts()),
              typename = typename std::enable_if<!std::is_same<Src, compressed2D>::value>::type>
    compressed2D& operator=(const Src& src)
    {
        nrows_ = src.num_rows();
        ncols_ = src.num_cols();
        assign_csr(src.host_starts(), src.host_indices(),
                   std::vector<Value>(src.host_values().begin(), src.host_values().end()));
        return *this;
    }

  private:
    void release()
    {
        if (h_) m4g_matrix_destroy(h_);
        h_ = nullptr;
    }
    void upload()
    {
        release();
        int st = m4g_matrix_create(&h_, int64_t(nrows_), int64_t(ncols_), int64_t(indices_.size()), starts_.data(),
                                   indices_.data(), data_.data(), M4G_IDX_I64, detail::value_code<Value>::code);
        detail::throw_status(st, "compressed2D");
    }

    size_type nrows_ = 0, ncols_ = 0;
    std::vector<int64_t> starts_{0}, indices_;
    std::vector<Value> data_;
    m4g_matrix h_ = nullptr;
};

template <typename V> inline std::size_t num_rows(const compressed2D<V>& A) { return A.num_rows(); }
template <typename V> inline std::size_t num_cols(const compressed2D<V>& A) { return A.num_cols(); }
template <typename V> inline std::size_t size(const compressed2D<V>& A) { return A.num_rows() * A.num_cols(); }

namespace detail {
// Host rows of a matrix kept beside its device format: what the inserter
// reads and rebuilds, and what element access and conversions use.
template <typename Value>
struct host_rows {
    std::size_t nrows_ = 0, ncols_ = 0;
    std::vector<int64_t> starts_{0}, indices_;
    std::vector<Value> values_;
    void reset(std::size_t m, std::size_t n)
    {
        nrows_ = m;
        ncols_ = n;
        starts_.assign(m + 1, 0);
        indices_.clear();
        values_.clear();
    }
    Value at(std::size_t r, std::size_t c) const
    {
        auto b = indices_.begin() + starts_[r], e = indices_.begin() + starts_[r + 1];
        auto p = std::lower_bound(b, e, int64_t(c));
        return (p != e && *p == int64_t(c)) ? values_[p - indices_.begin()] : Value(0);
    }
};
}  // namespace detail

// ell_matrix<Value> (mtl/matrix/ell_matrix.hpp:31-177): row-major ELLPACK with
// the reference's slot layout, held on the device (mtl4gpu_formats.h).
template <typename Value = double>
class ell_matrix {
    static_assert(std::is_same<Value, double>::value, "ell_matrix holds double values on the device");

  public:
    typedef Value value_type;
    typedef std::size_t size_type;
    static const unsigned alignment = 32;

    ell_matrix() = default;
    ell_matrix(size_type m, size_type n) { rows_.reset(m, n); }
    ell_matrix(const ell_matrix& o) : rows_(o.rows_) { upload(); }
    template <typename Src, typename = decltype(std::declval<const Src&>().host_starts()),
              typename = typename std::enable_if<!std::is_same<Src, ell_matrix>::value>::type>
    explicit ell_matrix(const Src& src)
    {
        *this = src;
    }
    ell_matrix& operator=(const ell_matrix& o)
    {
        if (this != &o) {
            rows_ = o.rows_;
            upload();
        }
        return *this;
    }
    template <typename Src, typename = decltype(std::declval<const Src&>().host_starts()),
              typename = typename std::enable_if<!std::is_same<Src, ell_matrix>::value>::type>
    ell_matrix& operator=(const Src& src)
    {
        rows_.reset(src.num_rows(), src.num_cols());
        assign_csr(src.host_starts(), src.host_indices(),
                   std::vector<Value>(src.host_values().begin(), src.host_values().end()));
        return *this;
    }
    ~ell_matrix() { release(); }

    void change_dim(size_type m, size_type n)
    {
        if (m != num_rows() || n != num_cols()) {
            rows_.reset(m, n);
            upload();
        }
    }
    void make_empty()
    {
        rows_.reset(num_rows(), num_cols());
        upload();
    }
    void assign_csr(std::vector<int64_t> S, std::vector<int64_t> I, std::vector<Value> D)
    {
        rows_.starts_ = std::move(S);
        rows_.indices_ = std::move(I);
        rows_.values_ = std::move(D);
        upload();
    }

    size_type num_rows() const { return rows_.nrows_; }
    size_type num_cols() const { return rows_.ncols_; }
    size_type nnz() const { return rows_.indices_.size(); }
    size_type slots() const { return slots_; }
    size_type stride() const { return stride_; }
    // the slot layout (ref_minor / ref_data, ell_matrix.hpp:77-80)
    const std::vector<size_type>& ref_minor() const { return minor_; }
    const std::vector<value_type>& ref_data() const { return data_; }
    Value operator()(size_type r, size_type c) const { return rows_.at(r, c); }
    struct row_proxy {
        const ell_matrix& A;
        size_type r;
        Value operator[](size_type c) const { return A(r, c); }
    };
    row_proxy operator[](size_type r) const { return {*this, r}; }

    const std::vector<int64_t>& host_starts() const { return rows_.starts_; }
    const std::vector<int64_t>& host_indices() const { return rows_.indices_; }
    const std::vector<Value>& host_values() const { return rows_.values_; }
    m4g_ell handle() const { return h_; }
    int mult_device(m4g_vector v, m4g_vector w, int mode) const { return m4g_ell_spmv(h_, v, w, mode); }

  private:
    void release()
    {
        if (h_) m4g_ell_destroy(h_);
        h_ = nullptr;
    }
    void upload()
    {
        release();
        mtl::detail::throw_status(m4g_ell_create(&h_, int64_t(num_rows()), int64_t(num_cols()), rows_.starts_.data(),
                                                 rows_.indices_.data(), rows_.values_.data()),
                                  "ell_matrix");
        int64_t sl = 0, st = 0;
        m4g_ell_shape(h_, nullptr, nullptr, &sl, &st, nullptr);
        slots_ = size_type(sl);
        stride_ = size_type(st);
        std::vector<int64_t> idx(slots_ * stride_);
        data_.assign(slots_ * stride_, Value(0));
        mtl::detail::throw_status(m4g_ell_download(h_, idx.data(), data_.data()), "ell_matrix");
        minor_.assign(idx.begin(), idx.end());
    }

    detail::host_rows<Value> rows_;
    size_type slots_ = 0, stride_ = 0;
    std::vector<size_type> minor_;
    std::vector<value_type> data_;
    m4g_ell h_ = nullptr;
};

// sparse_banded<Value> (mtl/matrix/sparse_banded.hpp:38-258): the stored
// diagonals of the matrix, held band-major on the device (mtl4gpu_formats.h).
template <typename Value = double>
class sparse_banded {
    static_assert(std::is_same<Value, double>::value, "sparse_banded holds double values on the device");

  public:
    typedef Value value_type;
    typedef std::size_t size_type;
    typedef long band_size_type;

    explicit sparse_banded(size_type m = 0, size_type n = 0) { rows_.reset(m, n); }
    sparse_banded(const sparse_banded& o) : rows_(o.rows_) { upload(); }
    template <typename Src, typename = decltype(std::declval<const Src&>().host_starts()),
              typename = typename std::enable_if<!std::is_same<Src, sparse_banded>::value>::type>
    explicit sparse_banded(const Src& src)
    {
        *this = src;
    }
    sparse_banded& operator=(const sparse_banded& o)
    {
        if (this != &o) {
            rows_ = o.rows_;
            upload();
        }
        return *this;
    }
    template <typename Src, typename = decltype(std::declval<const Src&>().host_starts()),
              typename = typename std::enable_if<!std::is_same<Src, sparse_banded>::value>::type>
    sparse_banded& operator=(const Src& src)
    {
        rows_.reset(src.num_rows(), src.num_cols());
        assign_csr(src.host_starts(), src.host_indices(),
                   std::vector<Value>(src.host_values().begin(), src.host_values().end()));
        return *this;
    }
    ~sparse_banded() { release(); }

    void change_dim(size_type m, size_type n)
    {
        rows_.reset(m, n);
        upload();
    }
    void make_empty()
    {
        rows_.reset(num_rows(), num_cols());
        upload();
    }
    void assign_csr(std::vector<int64_t> S, std::vector<int64_t> I, std::vector<Value> D)
    {
        rows_.starts_ = std::move(S);
        rows_.indices_ = std::move(I);
        rows_.values_ = std::move(D);
        upload();
    }

    size_type num_rows() const { return rows_.nrows_; }
    size_type num_cols() const { return rows_.ncols_; }
    // bands times rows, the stored entries (sparse_banded.hpp:125)
    size_type nnz() const { return bands_.size() * num_rows(); }
    const std::vector<band_size_type>& ref_bands() const { return bands_; }
    const value_type* ref_data() const { return data_.data(); }
    Value operator()(size_type r, size_type c) const { return rows_.at(r, c); }
    struct row_proxy {
        const sparse_banded& A;
        size_type r;
        Value operator[](size_type c) const { return A(r, c); }
    };
    row_proxy operator[](size_type r) const { return {*this, r}; }

    const std::vector<int64_t>& host_starts() const { return rows_.starts_; }
    const std::vector<int64_t>& host_indices() const { return rows_.indices_; }
    const std::vector<Value>& host_values() const { return rows_.values_; }
    m4g_banded handle() const { return h_; }
    int mult_device(m4g_vector v, m4g_vector w, int mode) const { return m4g_banded_spmv(h_, v, w, mode); }

  private:
    void release()
    {
        if (h_) m4g_banded_destroy(h_);
        h_ = nullptr;
    }
    void upload()
    {
        release();
        mtl::detail::throw_status(m4g_banded_from_csr(&h_, int64_t(num_rows()), int64_t(num_cols()),
                                                      rows_.starts_.data(), rows_.indices_.data(),
                                                      rows_.values_.data()),
                                  "sparse_banded");
        int64_t nb = 0;
        m4g_banded_shape(h_, nullptr, nullptr, &nb);
        std::vector<int64_t> b(static_cast<size_t>(nb));
        data_.assign(size_t(nb) * num_rows(), Value(0));
        mtl::detail::throw_status(m4g_banded_download(h_, b.data(), data_.data()), "sparse_banded");
        bands_.assign(b.begin(), b.end());
    }

    detail::host_rows<Value> rows_;
    std::vector<band_size_type> bands_;
    std::vector<value_type> data_;
    m4g_banded h_ = nullptr;
};

template <typename V> inline std::size_t num_rows(const ell_matrix<V>& A) { return A.num_rows(); }
template <typename V> inline std::size_t num_cols(const ell_matrix<V>& A) { return A.num_cols(); }
template <typename V> inline std::size_t size(const ell_matrix<V>& A) { return A.num_rows() * A.num_cols(); }
template <typename V> inline std::size_t num_rows(const sparse_banded<V>& A) { return A.num_rows(); }
template <typename V> inline std::size_t num_cols(const sparse_banded<V>& A) { return A.num_cols(); }
template <typename V> inline std::size_t size(const sparse_banded<V>& A) { return A.num_rows() * A.num_cols(); }

template <typename V>
inline mat_cvec_times_expr<compressed2D<V>, vec::dense_vector<V>> operator*(const compressed2D<V>& A,
                                                                            const vec::dense_vector<V>& v)
{
    return {A, v};
}

template <typename V>
inline mat_cvec_times_expr<ell_matrix<V>, vec::dense_vector<V>> operator*(const ell_matrix<V>& A,
                                                                          const vec::dense_vector<V>& v)
{
    return {A, v};
}
template <typename V>
inline mat_cvec_times_expr<sparse_banded<V>, vec::dense_vector<V>> operator*(const sparse_banded<V>& A,
                                                                            const vec::dense_vector<V>& v)
{
    return {A, v};
}

template <typename V>
inline residual_expr<vec::dense_vector<V>, compressed2D<V>> operator-(
    const vec::dense_vector<V>& b, const mat_cvec_times_expr<compressed2D<V>, vec::dense_vector<V>>& Ax)
{
    return {b, Ax};
}

template <typename V>
inline product_minus_expr<vec::dense_vector<V>, compressed2D<V>> operator-(
    const mat_cvec_times_expr<compressed2D<V>, vec::dense_vector<V>>& Ax, const vec::dense_vector<V>& b)
{
    return {Ax, b};
}

// update policies of the inserter (mtl/operation/update.hpp)
template <typename V> struct update_store { static V apply(V, V y) { return y; } };
template <typename V> struct update_plus { static V apply(V x, V y) { return x + y; } };

// inserter<Matrix, Updater>: collects entries, keeps rows sorted and merges
// duplicates with the updater (compressed2D.hpp:662-1151); the matrix is
// rebuilt when the inserter is destroyed.  Existing entries are kept.
template <typename Matrix, typename Updater = update_store<typename Matrix::value_type>>
class inserter {
    typedef typename Matrix::value_type V;

  public:
    explicit inserter(Matrix& A, std::size_t = 5) : A_(A)
    {
        for (std::size_t r = 0; r < num_rows(A); r++)
            for (int64_t p = A.host_starts()[r]; p < A.host_starts()[r + 1]; p++)
                rows_[{int64_t(r), A.host_indices()[p]}] = A.host_values()[p];
    }
    ~inserter()
    {
        std::vector<int64_t> S(num_rows(A_) + 1, 0), I;
        std::vector<V> D;
        for (auto& kv : rows_) S[kv.first.first + 1]++;
        for (std::size_t r = 0; r < num_rows(A_); r++) S[r + 1] += S[r];
        for (auto& kv : rows_) {
            I.push_back(kv.first.second);
            D.push_back(kv.second);
        }
        A_.assign_csr(std::move(S), std::move(I), std::move(D));
    }
    struct entry {
        inserter& ins;
        int64_t r, c;
        entry& operator<<(V v)
        {
            auto key = std::make_pair(r, c);
            auto it = ins.rows_.find(key);
            if (it == ins.rows_.end()) ins.rows_[key] = v;
            else it->second = Updater::apply(it->second, v);
            return *this;
        }
    };
    struct row {
        inserter& ins;
        int64_t r;
        entry operator[](std::size_t c) { return {ins, r, int64_t(c)}; }
    };
    entry operator()(std::size_t r, std::size_t c) { return {*this, int64_t(r), int64_t(c)}; }
    row operator[](std::size_t r) { return {*this, int64_t(r)}; }

  private:
    Matrix& A_;
    std::map<std::pair<int64_t, int64_t>, V> rows_;
};

// laplacian_setup(A, m, n) (laplacian_setup.hpp:25-44)
template <typename V>
inline void laplacian_setup(compressed2D<V>& A, unsigned m, unsigned n)
{
    const int64_t nnz = m4g_gen_laplacian2d(m, n, nullptr, nullptr, nullptr);
    std::vector<int64_t> S(std::size_t(m) * n + 1), I(nnz);
    std::vector<double> D(nnz);
    m4g_gen_laplacian2d(m, n, S.data(), I.data(), D.data());
    A.change_dim(std::size_t(m) * n, std::size_t(m) * n);
    A.assign_csr(std::move(S), std::move(I), std::vector<V>(D.begin(), D.end()));
}

// laplacian_setup for the alternative formats: the same rows
template <typename V>
inline void laplacian_setup(ell_matrix<V>& A, unsigned m, unsigned n)
{
    compressed2D<V> C;
    laplacian_setup(C, m, n);
    A = C;
}
template <typename V>
inline void laplacian_setup(sparse_banded<V>& A, unsigned m, unsigned n)
{
    compressed2D<V> C;
    laplacian_setup(C, m, n);
    A = C;
}

// poisson2D_dirichlet (mtl/matrix/poisson2D_dirichlet.hpp:24-102): the
// 5-point Laplacian of an m x n grid as a matrix-free operator, applied on
// the device (m4g_poisson2d_dirichlet_mult); operator* gives the
// mat_cvec_multiplier, so w = A * v, w += A * v and the solvers work on it.
struct poisson2D_dirichlet {
    typedef double value_type;
    typedef std::size_t size_type;
    poisson2D_dirichlet(int m, int n) : m(m), n(n), s(m * n) {}
    template <typename VectorIn, typename VectorOut, typename Assign>
    void mult(const VectorIn& v, VectorOut& w, Assign) const
    {
        MTL_THROW_IF(int(size(v)) != s, incompatible_size());
        if (size(w) == 0) w.change_resource(std::size_t(s));
        MTL_THROW_IF(size(v) != size(w), incompatible_size());
        mtl::detail::throw_status(
            m4g_poisson2d_dirichlet_mult(m, n, v.dev(), w.dev(), mtl::detail::assign_code<Assign>::value),
            "poisson2D_dirichlet");
        w.device_modified();
    }
    template <typename VectorIn>
    vec::mat_cvec_multiplier<poisson2D_dirichlet, VectorIn> operator*(const VectorIn& v) const
    {
        return vec::mat_cvec_multiplier<poisson2D_dirichlet, VectorIn>(*this, v);
    }
    int m, n, s;
};
inline std::size_t size(const poisson2D_dirichlet& A) { return std::size_t(A.s) * std::size_t(A.s); }
inline std::size_t num_rows(const poisson2D_dirichlet& A) { return std::size_t(A.s); }
inline std::size_t num_cols(const poisson2D_dirichlet& A) { return std::size_t(A.s); }

// trans(A): transposed copy (rows sorted)
template <typename V>
inline compressed2D<V> trans(const compressed2D<V>& A)
{
    const std::size_t m = num_rows(A), n = num_cols(A);
    std::vector<int64_t> S(n + 1, 0), I(A.nnz());
    std::vector<V> D(A.nnz());
    for (auto c : A.ref_minor()) S[c + 1]++;
    for (std::size_t r = 0; r < n; r++) S[r + 1] += S[r];
    std::vector<int64_t> fill(S.begin(), S.end() - 1);
    for (std::size_t r = 0; r < m; r++)
        for (int64_t p = A.ref_major()[r]; p < A.ref_major()[r + 1]; p++) {
            int64_t d = fill[A.ref_minor()[p]]++;
            I[d] = int64_t(r);
            D[d] = A.data()[p];
        }
    compressed2D<V> T(n, m);
    T.assign_csr(std::move(S), std::move(I), std::move(D));
    return T;
}

template <typename V>
inline std::ostream& operator<<(std::ostream& os, const compressed2D<V>& A)
{
    for (std::size_t r = 0; r < num_rows(A); r++) {
        os << "[";
        for (std::size_t c = 0; c < num_cols(A); c++) os << A(r, c) << (c + 1 < num_cols(A) ? " " : "");
        os << "]\n";
    }
    return os;
}

}  // namespace mat

using mat::compressed2D;
using mat::poisson2D_dirichlet;
using mat::ell_matrix;
using mat::sparse_banded;
using mat::inserter;
using mat::laplacian_setup;
using mat::num_cols;
using mat::num_rows;
using mat::trans;
using vec::size;
using vec::one_norm;

template <typename T> struct Collection { typedef typename T::value_type value_type; typedef std::size_t size_type; };

// resource(x): size of the temporaries (mtl/operation/resource.hpp:38-44)
template <typename V> inline std::size_t resource(const vec::dense_vector<V>& v) { return v.size(); }

namespace io {
// matrix_market_istream(path) >> A (matrix_market.hpp:44-217) through m4g_mm_read
class matrix_market_istream {
  public:
    explicit matrix_market_istream(const char* p) : path_(p) {}
    explicit matrix_market_istream(const std::string& p) : path_(p) {}
    template <typename V>
    matrix_market_istream& operator>>(mat::compressed2D<V>& A)
    {
        int64_t nr = 0, nc = 0, nz = 0;
        detail::throw_status(m4g_mm_read(path_.c_str(), &nr, &nc, &nz, nullptr, nullptr, nullptr),
                             "matrix_market_istream");
        std::vector<int64_t> S(size_t(nr) + 1), I(size_t(std::max<int64_t>(nz, 1)));
        std::vector<double> D(size_t(std::max<int64_t>(nz, 1)));
        detail::throw_status(m4g_mm_read(path_.c_str(), &nr, &nc, &nz, S.data(), I.data(), D.data()),
                             "matrix_market_istream");
        I.resize(size_t(nz));
        A.change_dim(std::size_t(nr), std::size_t(nc));
        A.assign_csr(std::move(S), std::move(I), std::vector<V>(D.begin(), D.begin() + nz));
        return *this;
    }
    void close() {}

  private:
    std::string path_;
};

// matrix_market_ostream(path) << A (matrix_market.hpp:219-335) through m4g_mm_write
class matrix_market_ostream {
  public:
    explicit matrix_market_ostream(const char* p) : path_(p) {}
    explicit matrix_market_ostream(const std::string& p) : path_(p) {}
    template <typename V>
    matrix_market_ostream& operator<<(const mat::compressed2D<V>& A)
    {
        std::vector<double> D(A.data().begin(), A.data().end());
        detail::throw_status(m4g_mm_write(path_.c_str(), int64_t(A.num_rows()), int64_t(A.num_cols()),
                                          A.ref_major().data(), A.ref_minor().data(), D.data()),
                             "matrix_market_ostream");
        return *this;
    }
    void close() {}

  private:
    std::string path_;
};

template <typename Matrix>
void matrix_market::read_into(Matrix& A) const
{
    matrix_market_istream(file_name_) >> A;
}
}  // namespace io

}  // namespace mtl

// ===========================================================================
// ITL
// ===========================================================================
namespace itl {

// ---------------------------------------------------------------------------
// iteration control (itl/iteration/basic_iteration.hpp:24-165)
// ---------------------------------------------------------------------------
template <class Real>
class basic_iteration {
  public:
    typedef basic_iteration self;
    typedef Real real;

    template <class Vector>
    basic_iteration(const Vector& r0, int max_iter_, Real t, Real a = Real(0))
        : error(0), i(0), my_norm_r0(std::abs(mtl::two_norm(r0))), max_iter(max_iter_), rtol_(t), atol_(a)
    {}
    basic_iteration(Real nb, int max_iter_, Real t, Real a = Real(0))
        : error(0), i(0), my_norm_r0(nb), max_iter(max_iter_), rtol_(t), atol_(a)
    {}
    virtual ~basic_iteration() {}

    bool check_max()
    {
        if (i >= max_iter) error = 1, is_finished = true, err_msg = "Too many iterations.";
        return is_finished;
    }
    bool finished(const Real& r)
    {
        if (converged(r)) return is_finished = true;
        return check_max();
    }
    template <class Vector>
    bool finished(const Vector& r) { return finished(Real(mtl::two_norm(r))); }
    bool finished() const { return is_finished; }

    self& operator++() { ++i; return *this; }
    self& operator+=(int n) { i += n; return *this; }
    bool first() const { return i <= 1; }
    virtual operator int() const { return error; }
    virtual int error_code() const { return error; }
    bool is_converged() const { return is_finished && error == 0; }
    int iterations() const { return i; }
    int max_iterations() const { return max_iter; }
    void set_max_iterations(int m) { max_iter = m; }
    void restart() { i = 0; }
    Real resid() const { return resid_; }
    Real relresid() const { return resid_ / my_norm_r0; }
    Real normb() const { return my_norm_r0; }
    Real norm_r0() const { return my_norm_r0; }
    void set_norm_r0(Real r0) { my_norm_r0 = r0; }
    Real tol() const { return rtol_; }
    Real atol() const { return atol_; }
    int fail(int err_code) { error = err_code; return error_code(); }
    int fail(int err_code, const std::string& msg) { error = err_code; err_msg = msg; return error_code(); }
    void set(Real v) { my_norm_r0 = v; }
    void set_quite(bool q) { my_quite = q; }
    bool is_quite() const { return my_quite; }
    void suppress_resume(bool s) { my_suppress = s; }
    bool resume_suppressed() const { return my_suppress; }

    // -- exchange with the device solver --------------------------------
    void to_c(m4g_iteration& c, std::vector<double>& hist) const
    {
        c.max_iter = max_iter;
        c.rtol = double(rtol_);
        c.atol = double(atol_);
        c.norm_r0 = double(my_norm_r0);
        c.iterations = i;
        c.error = error;
        c.resid = double(resid_);
        hist.assign(std::size_t(std::max(max_iter, i) + 2), std::nan(""));
        c.history = hist.data();
        c.history_capacity = int(hist.size());
        c.interrupt = nullptr;
    }
    // applies the device solve's outcome; logs the residuals checked in between
    virtual void from_c(const m4g_iteration& c, const std::vector<double>& hist)
    {
        i = c.iterations;
        resid_ = Real(c.resid);
        error = c.error;
        is_finished = true;
        (void)hist;
    }

  protected:
    bool converged(const Real& r)
    {
        resid_ = r;
        if (my_norm_r0 == 0) return resid_ <= atol_;
        return resid_ <= rtol_ * my_norm_r0 || resid_ <= atol_;
    }
    int error, i;
    Real my_norm_r0;
    int max_iter;
    Real rtol_, atol_, resid_ = Real(0);
    std::string err_msg;
    bool is_finished = false, my_quite = false, my_suppress = false;
};

// cyclic_iteration (itl/iteration/cyclic_iteration.hpp:23-101): prints
// "iteration i: resid r" every `cycle` iterations and a final summary.  The
// device solver reports the residuals of every finished() test at the end of
// the solve; they are printed then, in order.
template <class Real, class OStream = std::ostream>
class cyclic_iteration : public basic_iteration<Real> {
    typedef basic_iteration<Real> super;

  public:
    template <class Vector>
    cyclic_iteration(const Vector& r0, int max_iter_, Real tol_, Real atol_ = Real(0), int cycle_ = 100,
                     OStream& out = std::cout)
        : super(r0, max_iter_, tol_, atol_), cycle(cycle_), out(out)
    {}
    cyclic_iteration(Real r0, int max_iter_, Real tol_, Real atol_ = Real(0), int cycle_ = 100,
                     OStream& out = std::cout)
        : super(r0, max_iter_, tol_, atol_), cycle(cycle_), out(out)
    {}
    cyclic_iteration& operator++() { ++this->i; return *this; }
    operator int() const override { return error_code(); }
    void set_multi_print(bool m) { multi_print = m; }
    bool is_multi_print() const { return multi_print; }
    void set_cycle(int c) { cycle = c; }
    int error_code() const override
    {
        if (!this->my_suppress)
            out << "finished! error code = " << this->error << '\n'
                << this->iterations() << " iterations\n"
                << this->resid() << " is actual final residual. \n"
                << this->relresid() << " is actual relative tolerance achieved. \n"
                << "Relative tol: " << this->rtol_ << "  Absolute tol: " << this->atol_ << '\n'
                << "Convergence:  " << std::pow(this->relresid(), 1.0 / double(this->iterations())) << std::endl;
        return this->error;
    }
    void from_c(const m4g_iteration& c, const std::vector<double>& hist) override
    {
        const int i0 = this->i;
        super::from_c(c, hist);
        if (this->my_quite) return;
        for (int k = i0; k <= c.iterations && k < int(hist.size()); k++)
            if (k % cycle == 0 && !std::isnan(hist[k]) && (multi_print || k != last_print)) {
                out << "iteration " << k << ": resid " << hist[k] << std::endl;
                last_print = k;
            }
    }

  protected:
    int cycle, last_print = -1;
    bool multi_print = false;
    OStream& out;
};

template <class Real, class OStream = std::ostream>
class noisy_iteration : public cyclic_iteration<Real, OStream> {   // noisy_iteration.hpp:21-30
  public:
    template <class Vector>
    noisy_iteration(const Vector& r0, int max_iter_, Real tol_, Real atol_ = Real(0), OStream& out = std::cout)
        : cyclic_iteration<Real, OStream>(r0, max_iter_, tol_, atol_, 1, out)
    {}
};

// default_iteration (default_iteration.hpp:22-28): 100 iterations, 1e-10, quiet
template <class Real>
inline cyclic_iteration<Real> default_iteration()
{
    cyclic_iteration<Real> it(Real(0), 100, Real(1e-10), Real(0));
    it.set_quite(true);
    return it;
}

// ---------------------------------------------------------------------------
// preconditioners (itl/pc/)
// ---------------------------------------------------------------------------
namespace pc {

// solver<PC, Vector, adjoint>: lazy application, y = solve(P, x) (pc/solver.hpp:22-52)
template <typename PC, typename Vector, bool adjoint = false>
struct solver {
    typedef PC pc_type;
    const PC& P;
    const Vector& x;
    template <typename VectorOut>
    void assign_to(VectorOut& y) const
    {
        if (adjoint) P.adjoint_solve(x, y);
        else P.solve(x, y);
    }
};

class device_pc {
  public:
    // copies share the device factor (solver classes hold their preconditioner by value)
    m4g_pc handle() const { return h_.get(); }

    template <typename V>
    void solve(const mtl::dense_vector<V>& x, mtl::dense_vector<V>& y) const
    {
        y.checked_change_resource(x.size());
        if (!h_) mtl::detail::throw_status(m4g_vector_copy(y.dev(), x.dev()), "pc solve");   // identity of an operator
        else mtl::detail::throw_status(m4g_pc_solve(h_.get(), x.dev(), y.dev()), "pc solve");
        y.device_modified();
    }
    template <typename V>
    void adjoint_solve(const mtl::dense_vector<V>& x, mtl::dense_vector<V>& y) const
    {
        y.checked_change_resource(x.size());
        if (!h_) mtl::detail::throw_status(m4g_vector_copy(y.dev(), x.dev()), "pc adjoint_solve");
        else mtl::detail::throw_status(m4g_pc_adjoint_solve(h_.get(), x.dev(), y.dev()), "pc adjoint_solve");
        y.device_modified();
    }
    template <typename V>
    mtl::dense_vector<V> solve(const mtl::dense_vector<V>& x) const
    {
        mtl::dense_vector<V> y(x.size());
        solve(x, y);
        return y;
    }

  protected:
    template <typename Matrix>
    device_pc(const Matrix& A, int kind, int ftype)
    {
        if constexpr (mtl::detail::has_handle<Matrix>::value) {
            m4g_pc p = nullptr;
            mtl::detail::throw_status(m4g_pc_create(&p, A.handle(), kind, ftype), "preconditioner");
            h_.reset(p, [](m4g_pc q) { m4g_pc_destroy(q); });
        } else {
            // a matrix-free operator: only the identity needs no matrix entries
            MTL_THROW_IF(kind != M4G_PC_IDENTITY,
                         mtl::logic_error("this preconditioner needs the matrix entries (a compressed2D)"));
        }
    }
    template <typename V>
    mtl::mat::compressed2D<V> factor(int which) const
    {
        int64_t nnz = 0;
        mtl::detail::throw_status(m4g_pc_factor_nnz(h_.get(), which, &nnz), "factor");
        std::vector<int64_t> S(n_ + 1), I(nnz);
        std::vector<double> D(nnz);
        mtl::detail::throw_status(m4g_pc_download_factor(h_.get(), which, S.data(), I.data(), D.data()), "factor");
        mtl::mat::compressed2D<V> F(n_, n_);
        F.assign_csr(std::move(S), std::move(I), std::vector<V>(D.begin(), D.end()));
        return F;
    }
    std::shared_ptr<m4g_pc_s> h_;
    std::size_t n_ = 0;
};

template <typename Matrix, typename Value = typename mtl::Collection<Matrix>::value_type>
class identity : public device_pc {                 // pc/identity.hpp:24-71
  public:
    typedef Value value_type;
    explicit identity(const Matrix& A) : device_pc(A, M4G_PC_IDENTITY, M4G_F64) { n_ = num_rows(A); }
};

template <typename Matrix, typename Value = typename Matrix::value_type>
class diagonal : public device_pc {                 // pc/diagonal.hpp:28-90
  public:
    typedef Value value_type;
    explicit diagonal(const Matrix& A) : device_pc(A, M4G_PC_DIAGONAL, M4G_F64) { n_ = num_rows(A); }
};

template <typename Matrix, typename Value = typename Matrix::value_type>
class ic_0 : public device_pc {                     // pc/ic_0.hpp:39-194
  public:
    typedef Value value_type;
    typedef mtl::mat::compressed2D<Value> U_type;
    typedef U_type L_type;
    explicit ic_0(const Matrix& A) : device_pc(A, M4G_PC_IC0, mtl::detail::value_code<Value>::code)
    {
        n_ = num_rows(A);
    }
    L_type get_L() const { return factor<Value>(0); }
    U_type get_U() const { return factor<Value>(1); }
};

template <typename Matrix, typename Value = typename Matrix::value_type>
class ilu_0 : public device_pc {                    // pc/ilu_0.hpp:85-92, pc/ilu.hpp:33-133
  public:
    typedef Value value_type;
    typedef mtl::mat::compressed2D<Value> L_type;
    typedef mtl::mat::compressed2D<Value> U_type;
    explicit ilu_0(const Matrix& A) : device_pc(A, M4G_PC_ILU0, mtl::detail::value_code<Value>::code)
    {
        n_ = num_rows(A);
    }
    L_type get_L() const { return factor<Value>(0); }
    U_type get_U() const { return factor<Value>(1); }
};

template <typename PC, typename Vector>
inline solver<PC, Vector, false> solve(const PC& P, const Vector& x) { return {P, x}; }
template <typename PC, typename Vector>
inline solver<PC, Vector, true> adjoint_solve(const PC& P, const Vector& x) { return {P, x}; }

template <typename PC> inline bool is_identity(const PC&) { return false; }     // is_identity.hpp:21-27
template <typename M, typename V> inline bool is_identity(const identity<M, V>&) { return true; }
template <typename PC> struct static_is_identity : std::false_type {};
template <typename M, typename V> struct static_is_identity<identity<M, V>> : std::true_type {};

}  // namespace pc

// ---------------------------------------------------------------------------
// solvers
// ---------------------------------------------------------------------------
namespace detail {
template <typename Fn, typename Matrix, typename V, typename Iteration>
inline int run(Fn fn, const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b, m4g_pc P,
               Iteration& iter)
{
    m4g_iteration c;
    std::vector<double> hist;
    iter.to_c(c, hist);
    int st = fn(A.handle(), x.dev(), b.dev(), P, &c);
    x.device_modified();
    if (st == M4G_OK || st == M4G_MAX_ITER || st == M4G_BREAKDOWN || st == M4G_ERR_INTERRUPTED) {
        iter.from_c(c, hist);
        return iter;
    }
    mtl::detail::throw_status(st, "solver");
    return st;
}
}  // namespace detail

namespace detail {
// stands for "no preconditioner" in the operator CG (cg.hpp:39-72)
struct no_preconditioner {
    template <typename V> void solve(const V&, V&) const {}
};
}  // namespace detail
namespace pc {
inline bool is_identity(const detail::no_preconditioner&) { return true; }
}  // namespace pc
namespace detail {
// itl::cg for a linear operator without a device matrix (a matrix-free
// operator such as poisson2D_dirichlet or a user type with mult): the
// reference's steps on device vectors, q = A * p through the operator's
// multiplier.  Without a preconditioner cg.hpp:39-72, otherwise :75-112.
template <typename LinearOperator, typename V, typename Preconditioner, typename Iteration>
int cg_operator(const LinearOperator& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b,
                const Preconditioner* L, Iteration& iter)
{
    typedef mtl::dense_vector<V> Vector;
    typedef typename Iteration::real Real;
    auto ok = [](int st, const char* where) { mtl::detail::throw_status(st, where); };
    const std::size_t n = x.size();
    Vector p(n), q(n), r(n), z(n);
    q = A * x;                                                   // r = b - A*x
    ok(m4g_vector_copy(r.dev(), b.dev()), "cg");
    ok(m4g_axpy(-1.0, q.dev(), r.dev()), "cg");
    r.device_modified();
    double rho = 0, rho_1 = 0, rr = 0, alpha = 0, alpha_1 = 0;
    const bool plain = !L || pc::is_identity(*L);
    if (plain) ok(m4g_dot(r.dev(), r.dev(), &rho), "cg");
    else ok(m4g_dot(r.dev(), r.dev(), &rr), "cg");
    while (!iter.finished(Real(std::sqrt(std::abs(plain ? rho : rr))))) {
        ++iter;
        const Vector* src = &r;
        if (!plain) {
            L->solve(r, z);                                      // z = solve(L, r)
            ok(m4g_dot(r.dev(), z.dev(), &rho), "cg");
            src = &z;
        }
        if (iter.first()) ok(m4g_vector_copy(p.dev(), src->dev()), "cg");
        else ok(m4g_xpay(src->dev(), rho / rho_1, p.dev()), "cg");   // p = z + (rho / rho_1) * p
        p.device_modified();
        q = A * p;
        ok(m4g_dot(p.dev(), q.dev(), &alpha_1), "cg");
        alpha = rho / alpha_1;
        ok(m4g_axpy(alpha, p.dev(), x.dev()), "cg");             // x += alpha * p
        x.device_modified();
        rho_1 = rho;
        ok(m4g_axpy_unary_dot(alpha, q.dev(), r.dev(), plain ? &rho : &rr), "cg");   // r -= alpha * q; rr = r.r
        r.device_modified();
    }
    return iter;
}
}  // namespace detail

// itl::cg(A, x, b, P, iter) (cg.hpp:75-112); identity PC -> cg.hpp:39-72
template <typename Matrix, typename V, typename Preconditioner, typename Iteration>
inline int cg(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b, const Preconditioner& L,
              Iteration& iter)
{
    if constexpr (mtl::detail::has_handle<Matrix>::value) return detail::run(m4g_cg, A, x, b, L.handle(), iter);
    else return detail::cg_operator(A, x, b, &L, iter);
}
// no preconditioner (cg.hpp:39-72)
template <typename Matrix, typename V, typename Iteration>
inline int cg(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b, Iteration& iter)
{
    if constexpr (mtl::detail::has_handle<Matrix>::value) return detail::run(m4g_cg, A, x, b, nullptr, iter);
    else return detail::cg_operator(A, x, b, static_cast<const detail::no_preconditioner*>(nullptr), iter);
}
// right preconditioner ignored (cg.hpp:115-121)
template <typename Matrix, typename V, typename Preconditioner, typename RightPreconditioner, typename Iteration>
inline int cg(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b, const Preconditioner& L,
              const RightPreconditioner&, Iteration& iter)
{
    return cg(A, x, b, L, iter);
}
// itl::bicgstab(A, x, b, M, iter) (bicgstab.hpp:27-78)
template <typename Matrix, typename V, typename Preconditioner, typename Iteration>
inline int bicgstab(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b,
                    const Preconditioner& M, Iteration& iter)
{
    return detail::run(m4g_bicgstab, A, x, b, M.handle(), iter);
}

// itl::cgs(A, x, b, M, iter) (cgs.hpp:24-62)
template <typename Matrix, typename V, typename Preconditioner, typename Iteration>
inline int cgs(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b, const Preconditioner& M,
               Iteration& iter)
{
    return detail::run(m4g_cgs, A, x, b, M.handle(), iter);
}
// itl::bicg(A, x, b, M, iter) (bicg.hpp:27-68)
template <typename Matrix, typename V, typename Preconditioner, typename Iteration>
inline int bicg(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b, const Preconditioner& M,
                Iteration& iter)
{
    return detail::run(m4g_bicg, A, x, b, M.handle(), iter);
}
// itl::tfqmr(A, x, b, L, R, iter) (tfqmr.hpp:32-99)
template <typename Matrix, typename V, typename LeftPreconditioner, typename RightPreconditioner, typename Iteration>
inline int tfqmr(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b,
                 const LeftPreconditioner& L, const RightPreconditioner& R, Iteration& iter)
{
    const m4g_pc r = R.handle();
    auto fn = [r](m4g_matrix a, m4g_vector xx, m4g_vector bb, m4g_pc l, m4g_iteration* c) {
        return m4g_tfqmr(a, xx, bb, l, r, c);
    };
    return detail::run(fn, A, x, b, L.handle(), iter);
}

// itl::qmr(A, x, b, L, R, iter) (qmr.hpp:29-109)
template <typename Matrix, typename V, typename LeftPreconditioner, typename RightPreconditioner, typename Iteration>
inline int qmr(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b, const LeftPreconditioner& L,
               const RightPreconditioner& R, Iteration& iter)
{
    const m4g_pc r = R.handle();
    auto fn = [r](m4g_matrix a, m4g_vector xx, m4g_vector bb, m4g_pc l, m4g_iteration* c) {
        return m4g_qmr(a, xx, bb, l, r, c);
    };
    return detail::run(fn, A, x, b, L.handle(), iter);
}
// itl::gmres_full(A, x, b, L, R, iter) (gmres.hpp:37-119)
template <typename Matrix, typename V, typename LeftPreconditioner, typename RightPreconditioner, typename Iteration>
inline int gmres_full(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b,
                      const LeftPreconditioner& L, const RightPreconditioner& R, Iteration& iter)
{
    const m4g_pc r = R.handle();
    auto fn = [r](m4g_matrix a, m4g_vector xx, m4g_vector bb, m4g_pc l, m4g_iteration* c) {
        return m4g_gmres_full(a, xx, bb, l, r, c);
    };
    return detail::run(fn, A, x, b, L.handle(), iter);
}
// itl::gmres(A, x, b, L, R, iter, restart) (gmres.hpp:122-137)
template <typename Matrix, typename V, typename LeftPreconditioner, typename RightPreconditioner, typename Iteration>
inline int gmres(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b,
                 const LeftPreconditioner& L, const RightPreconditioner& R, Iteration& iter, std::size_t restart)
{
    const m4g_pc r = R.handle();
    const int64_t m = int64_t(restart);
    auto fn = [r, m](m4g_matrix a, m4g_vector xx, m4g_vector bb, m4g_pc l, m4g_iteration* c) {
        return m4g_gmres(a, xx, bb, l, r, c, m);
    };
    return detail::run(fn, A, x, b, L.handle(), iter);
}
// itl::bicgstab_2(A, x, b, L, iter) (bicgstab_2.hpp:35-130)
template <typename Matrix, typename V, typename Preconditioner, typename Iteration>
inline int bicgstab_2(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b,
                      const Preconditioner& L, Iteration& iter)
{
    return detail::run(m4g_bicgstab_2, A, x, b, L.handle(), iter);
}
// itl::bicgstab_ell(A, x, b, L, R, iter, l) (bicgstab_ell.hpp:37-135)
template <typename Matrix, typename V, typename LeftPreconditioner, typename RightPreconditioner, typename Iteration>
inline int bicgstab_ell(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b,
                        const LeftPreconditioner& L, const RightPreconditioner& R, Iteration& iter, std::size_t l)
{
    const m4g_pc r = R.handle();
    const int64_t ell = int64_t(l);
    auto fn = [r, ell](m4g_matrix a, m4g_vector xx, m4g_vector bb, m4g_pc lp, m4g_iteration* c) {
        return m4g_bicgstab_ell(a, xx, bb, lp, r, c, ell);
    };
    return detail::run(fn, A, x, b, L.handle(), iter);
}
// itl::idr_s(A, x, b, L, R, iter, s) (idr_s.hpp:35-103); L and R are ignored
template <typename Matrix, typename V, typename LeftPreconditioner, typename RightPreconditioner, typename Iteration>
inline int idr_s(const Matrix& A, mtl::dense_vector<V>& x, const mtl::dense_vector<V>& b,
                 const LeftPreconditioner& L, const RightPreconditioner& R, Iteration& iter, std::size_t s)
{
    const m4g_pc r = R.handle();
    const int64_t ss = int64_t(s);
    auto fn = [r, ss](m4g_matrix a, m4g_vector xx, m4g_vector bb, m4g_pc l, m4g_iteration* c) {
        return m4g_idr_s(a, xx, bb, l, r, c, ss);
    };
    return detail::run(fn, A, x, b, L.handle(), iter);
}

// base_solver (base_solver.hpp:24-84) and the solver classes (cg.hpp:125-179, bicgstab.hpp:84-112)
template <typename Solver, typename LinearOperator>
struct base_solver {
    typedef double magnitude_type;
    typedef cyclic_iteration<magnitude_type> iteration_type;

    base_solver(const LinearOperator& A, iteration_type iter = default_iteration<magnitude_type>())
        : A(A), my_iteration(iter), level(1)
    {
        my_iteration.set_quite(true);
    }
    template <typename HilbertSpaceB, typename HilbertSpaceX>
    int step(HilbertSpaceX& x, const HilbertSpaceB& b) const
    {
        basic_iteration<double> iter(b, 1, 0, 0);
        return static_cast<Solver const*>(this)->solve(x, b, iter);
    }
    template <typename HilbertSpaceB, typename HilbertSpaceX>
    int operator()(HilbertSpaceX& x, const HilbertSpaceB& b)
    {
        my_iteration.restart();
        if (mtl::two_norm(x) == 0) my_iteration.set_norm_r0(mtl::two_norm(b));
        else {
            HilbertSpaceB r(b);
            r -= A * x;
            my_iteration.set_norm_r0(mtl::two_norm(r));
        }
        return static_cast<Solver const*>(this)->solve(x, b, my_iteration);
    }
    void set_log_level(unsigned l)
    {
        level = l;
        my_iteration.set_quite(level < 2);
        my_iteration.suppress_resume(level == 0);
    }
    unsigned log_level() const { return level; }
    iteration_type iteration() const { return my_iteration; }
    iteration_type& iteration_ref() { return my_iteration; }

  protected:
    const LinearOperator& A;
    iteration_type my_iteration;
    unsigned level;
};

template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class cg_solver : public base_solver<cg_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<cg_solver, LinearOperator> base;

  public:
    explicit cg_solver(const LinearOperator& A) : base(A), L(A)
    {
        if (!pc::static_is_identity<RightPreconditioner>::value) std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    cg_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L)
    {
        if (!pc::static_is_identity<RightPreconditioner>::value) std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return cg(this->A, x, b, L, iter);
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b) const
    {
        basic_iteration<double> iter(x, 1, 0, 0);
        return solve(x, b, iter);
    }

  private:
    Preconditioner L;
};

template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class bicgstab_solver
    : public base_solver<bicgstab_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<bicgstab_solver, LinearOperator> base;

  public:
    explicit bicgstab_solver(const LinearOperator& A) : base(A), L(A)
    {
        if (!pc::static_is_identity<RightPreconditioner>::value) std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return bicgstab(this->A, x, b, L, iter);
    }

  private:
    Preconditioner L;
};

// cgs_solver (cgs.hpp:66-97), bicg_solver (bicg.hpp:72-103): right preconditioner ignored
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class cgs_solver : public base_solver<cgs_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<cgs_solver, LinearOperator> base;

  public:
    explicit cgs_solver(const LinearOperator& A) : base(A), L(A)
    {
        if (!pc::static_is_identity<RightPreconditioner>::value) std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return cgs(this->A, x, b, L, iter);
    }

  private:
    Preconditioner L;
};

template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class bicg_solver
    : public base_solver<bicg_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<bicg_solver, LinearOperator> base;

  public:
    explicit bicg_solver(const LinearOperator& A) : base(A), L(A)
    {
        if (!pc::static_is_identity<RightPreconditioner>::value) std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return bicg(this->A, x, b, L, iter);
    }

  private:
    Preconditioner L;
};

// tfqmr_solver (tfqmr.hpp:103-131): left and right preconditioners
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class tfqmr_solver
    : public base_solver<tfqmr_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<tfqmr_solver, LinearOperator> base;

  public:
    explicit tfqmr_solver(const LinearOperator& A) : base(A), L(A), R(A) {}
    tfqmr_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L), R(A) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return tfqmr(this->A, x, b, L, R, iter);
    }

  private:
    Preconditioner L;
    RightPreconditioner R;
};

// qmr_solver (qmr.hpp:113-140)
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class qmr_solver : public base_solver<qmr_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<qmr_solver, LinearOperator> base;

  public:
    explicit qmr_solver(const LinearOperator& A) : base(A), L(A), R(A) {}
    qmr_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L), R(A) {}
    qmr_solver(const LinearOperator& A, const Preconditioner& L, const RightPreconditioner& R) : base(A), L(L), R(R) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return qmr(this->A, x, b, L, R, iter);
    }

  private:
    Preconditioner L;
    RightPreconditioner R;
};

// gmres_solver (gmres.hpp:141-171), default restart 8
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class gmres_solver
    : public base_solver<gmres_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<gmres_solver, LinearOperator> base;

  public:
    explicit gmres_solver(const LinearOperator& A, std::size_t restart = 8) : base(A), restart(restart), L(A), R(A) {}
    gmres_solver(const LinearOperator& A, std::size_t restart, const Preconditioner& L)
        : base(A), restart(restart), L(L), R(A) {}
    gmres_solver(const LinearOperator& A, std::size_t restart, const Preconditioner& L, const RightPreconditioner& R)
        : base(A), restart(restart), L(L), R(R) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return gmres(this->A, x, b, L, R, iter, restart);
    }

  private:
    std::size_t restart;
    Preconditioner L;
    RightPreconditioner R;
};

// bicgstab_2_solver (bicgstab_2.hpp:136-166): right preconditioner ignored
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class bicgstab_2_solver
    : public base_solver<bicgstab_2_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<bicgstab_2_solver, LinearOperator> base;

  public:
    explicit bicgstab_2_solver(const LinearOperator& A) : base(A), L(A)
    {
        if (!pc::static_is_identity<RightPreconditioner>::value) std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    bicgstab_2_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L)
    {
        if (!pc::static_is_identity<RightPreconditioner>::value) std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return bicgstab_2(this->A, x, b, L, iter);
    }

  private:
    Preconditioner L;
};

// bicgstab_ell_solver (bicgstab_ell.hpp:140-168), default l = 8
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class bicgstab_ell_solver
    : public base_solver<bicgstab_ell_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<bicgstab_ell_solver, LinearOperator> base;

  public:
    explicit bicgstab_ell_solver(const LinearOperator& A, std::size_t l = 8) : base(A), l(l), L(A), R(A) {}
    bicgstab_ell_solver(const LinearOperator& A, std::size_t l, const Preconditioner& L) : base(A), l(l), L(L), R(A) {}
    bicgstab_ell_solver(const LinearOperator& A, std::size_t l, const Preconditioner& L, const RightPreconditioner& R)
        : base(A), l(l), L(L), R(R) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return bicgstab_ell(this->A, x, b, L, R, iter, l);
    }

  private:
    std::size_t l;
    Preconditioner L;
    RightPreconditioner R;
};

// idr_s_solver (idr_s.hpp:107-136), default s = 8
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class idr_s_solver
    : public base_solver<idr_s_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<idr_s_solver, LinearOperator> base;

  public:
    explicit idr_s_solver(const LinearOperator& A, std::size_t s = 8) : base(A), s(s), L(A), R(A) {}
    idr_s_solver(const LinearOperator& A, std::size_t s, const Preconditioner& L) : base(A), s(s), L(L), R(A) {}
    idr_s_solver(const LinearOperator& A, std::size_t s, const Preconditioner& L, const RightPreconditioner& R)
        : base(A), s(s), L(L), R(R) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        return idr_s(this->A, x, b, L, R, iter, s);
    }

  private:
    std::size_t s;
    Preconditioner L;
    RightPreconditioner R;
};

// make_cg_solver(A, L) (cg.hpp:174-179)
template <typename LinearOperator, typename Preconditioner>
inline cg_solver<LinearOperator, Preconditioner, pc::identity<LinearOperator>> make_cg_solver(const LinearOperator& A,
                                                                                                const Preconditioner& L)
{
    return cg_solver<LinearOperator, Preconditioner, pc::identity<LinearOperator>>(A, L);
}

// interruptible_iteration (iteration/interruptible_iteration.hpp:21-52): a
// basic_iteration plus a flag another thread may raise.  Device solves poll
// the flag between iteration batches (the device-resident drivers) or at every
// residual test (the host-driven ones); an interrupted solve returns the
// wrapped iteration's error code, as the reference's finished() does.
class interruptible_iteration {
  public:
    typedef double real;
    interruptible_iteration(basic_iteration<double>& iter) : iter(&iter), interrupted(false) {}
    interruptible_iteration() : iter(nullptr), interrupted(false) {}
    void set_iter(basic_iteration<double>& new_iter) { iter = &new_iter; }
    template <typename T>
    bool finished(const T& r)
    {
        return iter->finished(r) || interrupted.load();
    }
    void interrupt()
    {
        flag_ = 1;
        interrupted = true;
    }
    bool is_interrupted() const { return interrupted.load(); }
    operator int() const { return *iter; }
    int error_code() const { return iter->error_code(); }
    interruptible_iteration& operator++()
    {
        ++*iter;
        return *this;
    }
    bool first() const { return iter->first(); }
    int iterations() const { return iter->iterations(); }

    void to_c(m4g_iteration& c, std::vector<double>& hist) const
    {
        iter->to_c(c, hist);
        c.interrupt = &flag_;
    }
    void from_c(const m4g_iteration& c, const std::vector<double>& hist)
    {
        m4g_iteration d = c;
        if (d.error == M4G_ERR_INTERRUPTED) d.error = iter->error_code();
        iter->from_c(d, hist);
    }

  private:
    basic_iteration<double>* iter;
    std::atomic<bool> interrupted;
    volatile int flag_ = 0;
};

// async_executor (krylov/async_executor.hpp:23-74): runs a solver on its own
// host thread with an interruptible iteration; the thread's library calls go
// to that thread's stream of the same device.
template <typename Solver>
class async_executor {
  public:
    async_executor(const Solver& solver) : my_solver(solver), my_iter(), my_thread() {}
    ~async_executor()
    {
        if (my_thread.joinable()) my_thread.join();
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    void start_solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        my_iter.set_iter(iter);
        my_thread = std::thread([this, &x, &b]() { my_solver.solve(x, b, my_iter); });
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const
    {
        start_solve(x, b, iter);
        return const_cast<async_executor*>(this)->wait();
    }
    template <typename HilbertSpaceX, typename HilbertSpaceB>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b) const
    {
        basic_iteration<double> iter(x, 1, 0, 0);
        return solve(x, b, iter);
    }
    int wait()
    {
        my_thread.join();
        return my_iter.error_code();
    }
    bool is_finished() const { return !my_thread.joinable(); }
    int interrupt()
    {
        my_iter.interrupt();
        return wait();
    }

  private:
    Solver my_solver;
    mutable interruptible_iteration my_iter;
    mutable std::thread my_thread;
};

template <typename Solver>
inline async_executor<Solver> make_async_executor(const Solver& solver)
{
    return async_executor<Solver>(solver);
}

}  // namespace itl

namespace mtl {
namespace detail {
// the library's random(P) counter seen as an int (mtl::seed<double>::counter)
struct seed_counter {
    operator int() const { return m4g_random_seed_counter(); }
    seed_counter& operator=(int c)
    {
        m4g_set_random_seed_counter(c);
        return *this;
    }
};
}  // namespace detail
// mtl::seed (operation/random.hpp:29-38): only its counter is needed, which
// the library keeps for idr_s's random(P)
template <typename T>
struct seed {
    static inline detail::seed_counter counter{};
};
}  // namespace mtl

#endif  // MTL4GPU_HPP
