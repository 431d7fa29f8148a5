// T-itl/interrupted_cg_test.cpp through the drop-in headers: a CG solve on the
// 300 x 300 Laplacian with an unreachable tolerance runs on its own thread
// (async_executor of make_cg_solver(A, P)) and is interrupted after 50 ms;
// then a normal solve through the same executor type converges.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <chrono>
#include <iostream>
#include <thread>

int main()
{
    const int size = 300, N = size * size;
    typedef mtl::compressed2D<double> matrix_type;
    matrix_type A(N, N);
    laplacian_setup(A, size, size);
    itl::pc::identity<matrix_type> P(A);
    mtl::dense_vector<double> x(N, 1.0), b(N);
    b = A * x;
    x = 0;
    const int iter_limit = 100000;
    itl::cyclic_iteration<double> iter(b, iter_limit, 0.0, 0.0, iter_limit);
    auto my_cg_solver = itl::make_cg_solver(A, P);
    itl::async_executor<decltype(my_cg_solver)> async_exec(my_cg_solver);
    async_exec.start_solve(x, b, iter);
    std::this_thread::sleep_for(std::chrono::milliseconds(50));
    const int code = async_exec.interrupt();
    std::cout << "interrupted after " << iter.iterations() << " iterations, code " << code << "\n";
    MTL_THROW_IF(iter.iterations() == iter_limit, mtl::runtime_error("Executor was not interrupted!"));
    MTL_THROW_IF(code != 0, mtl::runtime_error("an interrupted solve keeps the iteration's error code"));

    // a solve that finishes: the executor waits for it
    itl::pc::ic_0<matrix_type> IC(A);
    x = 0;
    itl::basic_iteration<double> it2(b, 2000, 1.e-8);
    auto exec2 = itl::make_async_executor(itl::make_cg_solver(A, IC));
    MTL_THROW_IF(exec2.solve(x, b, it2) != 0, mtl::runtime_error("async solve did not converge"));
    MTL_THROW_IF(!exec2.is_finished(), mtl::runtime_error("is_finished"));
    std::cout << "async solve: " << it2.iterations() << " iterations\n";
    return 0;
}
