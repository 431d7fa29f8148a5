// The CPU-runnable headline configuration through the drop-in call surface:
// 2D 5-point Poisson 64x64, b = A x_true (x_true ~ U[-1,1], seed 1),
// itl::cg + itl::pc::ic_0 to 1e-10.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <vector>

int main()
{
    const int size = 64, N = size * size;
    typedef mtl::compressed2D<double> matrix_type;
    matrix_type A(N, N);
    laplacian_setup(A, size, size);
    std::vector<double> xt(N);
    m4g_rng_uniform(1, 1, 0, N, -1.0, 1.0, xt.data());
    mtl::dense_vector<double> x_true(N), b(N), x(N);
    for (int i = 0; i < N; i++) x_true[i] = xt[i];
    b = A * x_true;
    x = 0;
    itl::pc::ic_0<matrix_type> P(A);
    itl::cyclic_iteration<double> iter(b, 1000, 1.e-10, 0.0, 50);
    int err = cg(A, x, b, P, iter);
    mtl::dense_vector<double> r(b - A * x);
    std::cout << iter.iterations() << " iterations, true relres " << two_norm(r) / two_norm(b) << "\n";
    MTL_THROW_IF(err != 0 || two_norm(r) > 1e-9 * two_norm(b), mtl::runtime_error("64x64 IC(0)-PCG failed"));
    return 0;
}
