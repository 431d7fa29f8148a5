// T-mtl/matrix_free_1_test.cpp and matrix_free_2_test.cpp (a user operator
// whose operator* returns the product vector), and examples/matrix_free_cg.cpp
// (the library's poisson2D_dirichlet with the multiplier protocol, solved by
// itl::cg with the identity preconditioner) through the drop-in headers.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <cassert>
#include <iostream>

struct user_poisson2D {   // matrix_free_2_test.cpp's operator, restated
    user_poisson2D(int m, int n) : m(m), n(n) {}
    template <typename Vector>
    Vector operator*(const Vector& v) const
    {
        assert(int(size(v)) == m * n);
        Vector w(m * n);
        for (int i = 1; i < m - 1; i++)
            for (int j = 1, k = i * n + j; j < n - 1; j++, k++) w[k] = 4 * v[k] - v[k - n] - v[k + n] - v[k - 1] - v[k + 1];
        for (int j = 1; j < n - 1; j++) w[j] = 4 * v[j] - v[j + n] - v[j - 1] - v[j + 1];
        for (int j = 1, k = (m - 1) * n + j; j < n - 1; j++, k++) w[k] = 4 * v[k] - v[k - n] - v[k - 1] - v[k + 1];
        for (int i = 1, k = n; i < m - 1; i++, k += n) w[k] = 4 * v[k] - v[k - n] - v[k + n] - v[k + 1];
        for (int i = 1, k = n + n - 1; i < m - 1; i++, k += n) w[k] = 4 * v[k] - v[k - n] - v[k + n] - v[k - 1];
        w[0] = 4 * v[0] - v[1] - v[n];
        w[n - 1] = 4 * v[n - 1] - v[n - 2] - v[2 * n - 1];
        w[(m - 1) * n] = 4 * v[(m - 1) * n] - v[(m - 2) * n] - v[(m - 1) * n + 1];
        w[m * n - 1] = 4 * v[m * n - 1] - v[m * n - 2] - v[m * n - n - 1];
        return w;
    }
    int m, n;
};
namespace mtl { namespace ashape {
template <> struct ashape_aux<user_poisson2D> { typedef nonscal type; };
}}

int main()
{
    typedef mtl::dense_vector<double> vt;
    {   // matrix_free_1/2: the operator's product equals the matrix's
        mtl::compressed2D<double> A0;
        laplacian_setup(A0, 4, 5);
        vt v(20);
        iota(v);
        vt w1(A0 * v);
        user_poisson2D A(4, 5);
        vt w2;
        w2 = A * v;
        vt d(w1 - w2);
        MTL_THROW_IF(one_norm(d) > 0.001, mtl::runtime_error("user operator: wrong result"));
        mtl::mat::poisson2D_dirichlet P2(4, 5);
        vt w3(20);
        w3 = P2 * v;
        vt d3(w1 - w3);
        MTL_THROW_IF(one_norm(d3) != 0.0, mtl::runtime_error("poisson2D_dirichlet: wrong result"));
        w3 += P2 * v;                     // the incrementer: 2 A v
        w3 -= P2 * v;
        vt d4(w1 - w3);
        MTL_THROW_IF(one_norm(d4) != 0.0, mtl::runtime_error("poisson2D_dirichlet: += / -="));
    }
    {   // examples/matrix_free_cg.cpp
        const int size = 10, N = size * size;
        typedef mtl::mat::poisson2D_dirichlet matrix_type;
        matrix_type A(size, size);
        itl::pc::identity<matrix_type> P(A);
        vt x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        itl::cyclic_iteration<double> iter(b, 100, 1.e-11, 0.0, 5);
        MTL_THROW_IF(cg(A, x, b, P, iter) != 0, mtl::runtime_error("matrix-free cg did not converge"));
        for (int i = 0; i < N; i++) MTL_THROW_IF(std::abs(x[i] - 1.0) > 1e-9, mtl::runtime_error("matrix-free cg x"));
        // the same operator on a user type: its product runs through the host mirror
        user_poisson2D U(size, size);
        vt y(N, 0.0);
        itl::basic_iteration<double> it2(b, 100, 1.e-11);
        MTL_THROW_IF(cg(U, y, b, it2) != 0, mtl::runtime_error("user-operator cg did not converge"));
        MTL_THROW_IF(it2.iterations() != iter.iterations(), mtl::runtime_error("iteration counts differ"));
    }
    std::cout << "matrix-free ok\n";
    return 0;
}
