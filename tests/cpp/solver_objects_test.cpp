// The solver classes (cg_solver, bicgstab_solver; cg.hpp:125-165,
// bicgstab.hpp:84-112): constructed from the matrix, called as s(x, b), their
// iteration object reachable through iteration_ref().  Each is run on three
// grids; the default iteration (default_iteration.hpp:22-28: 100 iterations,
// relative 1e-10) must converge and the true residual must match it.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include "check.hpp"

namespace {

using Mat = mtl::compressed2D<double>;

template <typename Solver>
void run(Checks& c, const char* name, int side)
{
    const int n = side * side;
    Mat K;
    laplacian_setup(K, side, side);
    mtl::dense_vector<double> ones(n, 1.0), rhs(n), sol(n, 0.0);
    rhs = K * ones;
    Solver s(K);
    s(sol, rhs);
    mtl::dense_vector<double> res(rhs - K * sol);
    const double rel = two_norm(res) / two_norm(rhs);
    std::printf("%s on %dx%d: %d iterations, converged %d, true relres %.2e\n", name, side, side,
                s.iteration_ref().iterations(), int(s.iteration_ref().is_converged()), rel);
    c.expect(s.iteration_ref().is_converged(), name);
    c.expect(rel <= 1e-7, "true residual of a converged solver object");
}

}  // namespace

int main()
{
    Checks c;
    for (int side : {3, 6, 11}) {
        run<itl::cg_solver<Mat, itl::pc::ic_0<Mat>>>(c, "cg_solver<ic_0>", side);
        run<itl::cg_solver<Mat>>(c, "cg_solver<identity>", side);
        run<itl::bicgstab_solver<Mat, itl::pc::ilu_0<Mat>>>(c, "bicgstab_solver<ilu_0>", side);
    }
    return c.status();
}
