// The rest of the ITL Krylov family through the drop-in headers.
//  * Free functions with preconditioners: cgs, bicg, tfqmr on the 4x4-grid
//    Laplacian from x = 0.01 to 1e-4 (the reference's ic_0_krylov_test,
//    T-itl/, expects each to return 0); qmr with diagonal L and R on the 8x8
//    grid to 1e-6 (qmr_test); IDR(4), whose shadow space consumes one seed
//    per column, on the 10x10 grid to 1e-8 (idr_s_test); BiCGStab(2) + IC(0)
//    on the 10x10 grid to 1e-6 (ic_0_bicgstab_2_test).
//  * GMRES(1) from x = 5 on the 2x2-grid Laplacian with b = ones: one
//    minimal-residual step is exact there (gmres_test's known answer).
//  * Every solver class, with and without ILU(0), on the 3x3 grid with
//    n + 1 iterations allowed (the use in solver_classes_test).
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <functional>
#include <string>
#include <vector>
#include "check.hpp"

namespace {

using Mat = mtl::compressed2D<double>;
using Vec = mtl::dense_vector<double>;

struct Grid {
    int n;
    Mat K;
    Vec rhs;
    explicit Grid(int side) : n(side * side), K(side * side, side * side), rhs(side * side)
    {
        laplacian_setup(K, side, side);
        Vec ones(n, 1.0);
        rhs = K * ones;
    }
    double relres(const Vec& x) const
    {
        Vec r(rhs - K * x);
        return two_norm(r) / two_norm(rhs);
    }
};

// runs `solve` from x0 and checks its return code and true residual
void expect_solve(Checks& c, const std::string& name, const Grid& G, double x0, double tol,
                  const std::function<int(Vec&, itl::cyclic_iteration<double>&)>& solve)
{
    Vec x(G.n, x0);
    itl::cyclic_iteration<double> ctl(G.rhs, 500, tol, 0.0, 500);
    const int code = solve(x, ctl);
    const double rel = G.relres(x);
    std::printf("%-24s code %d, %3d iterations, true relres %.2e\n", name.c_str(), code, ctl.iterations(), rel);
    c.expect(code == 0, (name + " returned 0").c_str());
    // x0 != 0: the iteration's norm_r0 is |b|, so the residual bound is relative to |b|
    c.expect(rel <= 10 * tol, (name + " true residual").c_str());
}

template <typename Solver>
void solver_class(Checks& c, const char* name)
{
    Grid G(3);
    Vec x(G.n, 0.0);
    Solver s(G.K);
    s.iteration_ref().set_max_iterations(G.n + 1);
    s.iteration_ref().suppress_resume(true);
    s(x, G.rhs);
    std::printf("%-40s %d iterations, relres %.2e\n", name, s.iteration().iterations(), G.relres(x));
    c.expect(s.iteration().iterations() <= G.n + 1, name);
}

}  // namespace

int main()
{
    Checks c;
    {
        Grid G(4);
        itl::pc::identity<Mat> none(G.K);
        itl::pc::ilu_0<Mat> lu(G.K);
        expect_solve(c, "bicg (identity)", G, 0.01, 1e-4, [&](Vec& x, auto& it) { return itl::bicg(G.K, x, G.rhs, none, it); });
        expect_solve(c, "cgs (ILU(0))", G, 0.01, 1e-4, [&](Vec& x, auto& it) { return itl::cgs(G.K, x, G.rhs, lu, it); });
        expect_solve(c, "tfqmr (ILU(0), identity)", G, 0.01, 1e-4,
                     [&](Vec& x, auto& it) { return itl::tfqmr(G.K, x, G.rhs, lu, none, it); });
    }
    {
        Grid G(8);
        itl::pc::diagonal<Mat> jac(G.K);
        expect_solve(c, "qmr (diagonal, diagonal)", G, 0.0, 1e-6,
                     [&](Vec& x, auto& it) { return itl::qmr(G.K, x, G.rhs, jac, jac, it); });
    }
    {
        Grid G(10);
        itl::pc::identity<Mat> none(G.K);
        itl::pc::ic_0<Mat> ic(G.K);
        mtl::seed<double>::counter = 0;
        expect_solve(c, "idr_s(4)", G, 0.0, 1e-8, [&](Vec& x, auto& it) { return itl::idr_s(G.K, x, G.rhs, none, none, it, 4); });
        c.expect(int(mtl::seed<double>::counter) == 4, "idr_s(4) draws one seed per shadow column");
        expect_solve(c, "bicgstab_2 (IC(0))", G, 0.0, 1e-6, [&](Vec& x, auto& it) { return itl::bicgstab_2(G.K, x, G.rhs, ic, it); });
    }
    {
        Mat K(4, 4);
        laplacian_setup(K, 2, 2);
        itl::pc::identity<Mat> none(K);
        Vec b(4, 1.0), x(4, 5.0);
        itl::basic_iteration<double> one_step(b, 1, 1e-8);
        itl::gmres(K, x, b, none, none, one_step, 1);
        Vec r(K * x - b);
        std::printf("gmres(1), one step from x = 5: |A x - b| = %.2e\n", two_norm(r));
        c.expect(two_norm(r) <= 5e-5, "one GMRES(1) step solves the 2x2-grid system");
    }
    using itl::pc::ilu_0;
    solver_class<itl::cgs_solver<Mat>>(c, "cgs_solver");
    solver_class<itl::cgs_solver<Mat, ilu_0<Mat>>>(c, "cgs_solver<ilu_0>");
    solver_class<itl::bicg_solver<Mat>>(c, "bicg_solver");
    solver_class<itl::tfqmr_solver<Mat>>(c, "tfqmr_solver");
    solver_class<itl::tfqmr_solver<Mat, ilu_0<Mat>>>(c, "tfqmr_solver<ilu_0>");
    solver_class<itl::bicgstab_2_solver<Mat>>(c, "bicgstab_2_solver");
    solver_class<itl::bicgstab_2_solver<Mat, ilu_0<Mat>>>(c, "bicgstab_2_solver<ilu_0>");
    solver_class<itl::bicgstab_ell_solver<Mat>>(c, "bicgstab_ell_solver");
    solver_class<itl::bicgstab_ell_solver<Mat, ilu_0<Mat>>>(c, "bicgstab_ell_solver<ilu_0>");
    solver_class<itl::gmres_solver<Mat>>(c, "gmres_solver");
    solver_class<itl::gmres_solver<Mat, ilu_0<Mat>>>(c, "gmres_solver<ilu_0>");
    solver_class<itl::gmres_solver<Mat, ilu_0<Mat>, ilu_0<Mat>>>(c, "gmres_solver<ilu_0, ilu_0>");
    solver_class<itl::qmr_solver<Mat>>(c, "qmr_solver");
    solver_class<itl::qmr_solver<Mat, ilu_0<Mat>>>(c, "qmr_solver<ilu_0>");
    solver_class<itl::idr_s_solver<Mat>>(c, "idr_s_solver");
    solver_class<itl::idr_s_solver<Mat, ilu_0<Mat>>>(c, "idr_s_solver<ilu_0>");
    return c.status();
}
