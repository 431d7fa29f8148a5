// Interrupting a solve that runs on its own host thread.  Known answer of the
// reference (T-itl/interrupted_cg_test.cpp): a CG solve on the 300x300-grid
// Laplacian whose tolerances cannot be met, started through an
// async_executor and interrupted, stops before its iteration limit, and the
// interrupt returns the iteration's error code (0: no error was recorded).
// Then an IC(0)-preconditioned solve through make_async_executor runs to
// convergence and the executor reports it finished.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <chrono>
#include <cstdio>
#include <thread>

int main()
{
    using Mat = mtl::compressed2D<double>;
    const int side = 300, n = side * side;
    Mat K(n, n);
    laplacian_setup(K, side, side);
    mtl::dense_vector<double> ones(n, 1.0), rhs(n), sol(n, 0.0);
    rhs = K * ones;

    // 1. a solve that would run for a long time, stopped from this thread
    const int never = 1000000;
    itl::pc::identity<Mat> none(K);
    itl::cyclic_iteration<double> endless(rhs, never, 0.0, 0.0, never);
    const auto plain_cg = itl::make_cg_solver(K, none);
    itl::async_executor<decltype(plain_cg)> runner(plain_cg);
    runner.start_solve(sol, rhs, endless);
    std::this_thread::sleep_for(std::chrono::milliseconds(40));
    const int stop_code = runner.interrupt();
    const int done = endless.iterations();
    std::printf("stopped after %d of %d iterations; interrupt returned %d\n", done, never, stop_code);
    if (done <= 0 || done >= never) {
        std::printf("the background solve did not stop on the interrupt\n");
        return 1;
    }
    if (stop_code != 0) {
        std::printf("an interrupted solve must keep the iteration's own error code\n");
        return 1;
    }

    // 2. a solve that converges; solve() waits for its thread
    itl::pc::ic_0<Mat> ic(K);
    sol = 0;
    itl::basic_iteration<double> ctl(rhs, 2000, 1.0e-8);
    auto waiter = itl::make_async_executor(itl::make_cg_solver(K, ic));
    const int code = waiter.solve(sol, rhs, ctl);
    std::printf("converging solve: code %d after %d iterations, finished %d\n", code, ctl.iterations(),
                int(waiter.is_finished()));
    if (code != 0 || !waiter.is_finished()) {
        std::printf("the waited-for solve did not converge\n");
        return 1;
    }
    return 0;
}
