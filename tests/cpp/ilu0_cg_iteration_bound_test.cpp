// ILU(0) as CG's preconditioner on small grid Laplacians.  Known answer of
// the reference (T-itl/ilu_0_cg_test.cpp:33): on the 3x3 grid, CG reaches a
// relative residual of 1e-6 from x = 0, b = A * ones within 4 iterations.
// For the larger grids only convergence within n iterations (the
// exact-arithmetic bound of CG) is required.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <cstdio>

namespace {

struct Outcome {
    int code, iterations;
    double relres;
};

Outcome ilu_cg(int side)
{
    using Mat = mtl::compressed2D<double>;
    const int n = side * side;
    Mat K;
    laplacian_setup(K, side, side);
    itl::pc::ilu_0<Mat> lu(K);
    mtl::dense_vector<double> ones(n, 1.0), rhs(n), sol(n, 0.0);
    rhs = K * ones;
    itl::cyclic_iteration<double> ctl(rhs, n, 1.0e-6, 0.0, n);
    Outcome o;
    o.code = itl::cg(K, sol, rhs, lu, ctl);
    o.iterations = ctl.iterations();
    mtl::dense_vector<double> res(rhs - K * sol);
    o.relres = two_norm(res) / two_norm(rhs);
    return o;
}

}  // namespace

int main()
{
    bool ok = true;
    for (int side : {3, 5, 8}) {
        const Outcome o = ilu_cg(side);
        const int bound = side == 3 ? 4 : side * side;
        std::printf("%dx%d grid: code %d, %d iterations (bound %d), true relres %.2e\n", side, side, o.code,
                    o.iterations, bound, o.relres);
        ok = ok && o.code == 0 && o.iterations <= bound && o.relres <= 2.0e-6;
    }
    if (!ok) {
        std::printf("ILU(0)-preconditioned CG missed its iteration bound\n");
        return 1;
    }
    return 0;
}
