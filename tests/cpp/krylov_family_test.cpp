// The remaining solvers of the ITL Krylov family through the drop-in headers:
// the solver classes of T-itl/solver_classes_test.cpp (3x3-grid Laplacian,
// b = A * ones, x = 0, max_iterations = N + 1), gmres_test.cpp's GMRES(1)
// known answer, and the runs of qmr_test.cpp (diagonal L and R),
// idr_s_test.cpp (IDR(4)) and ic_0_bicgstab_2_test.cpp (IC(0)), which must
// reach their tolerances.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <cmath>
#include <iostream>

typedef mtl::compressed2D<double> matrix_type;

template <typename Solver>
void run_class(const char* name)
{
    const int size = 3, N = size * size;
    matrix_type A(N, N);
    laplacian_setup(A, size, size);
    mtl::dense_vector<double> x(N, 1.0), b(N);
    b = A * x;
    x = 0;
    Solver s(A);
    s.iteration_ref().set_max_iterations(N + 1);
    s.iteration_ref().suppress_resume(true);
    s(x, b);
    std::cout << name << ": " << s.iteration().iterations() << " iterations\n";
}

int main()
{
    using itl::pc::ilu_0;
    run_class<itl::bicgstab_2_solver<matrix_type>>("BiCGStab(2)");
    run_class<itl::bicgstab_2_solver<matrix_type, ilu_0<matrix_type>>>("BiCGStab(2) with ILU_0");
    run_class<itl::bicgstab_ell_solver<matrix_type>>("BiCGStab(ell)");
    run_class<itl::bicgstab_ell_solver<matrix_type, ilu_0<matrix_type>>>("BiCGStab(ell) with ILU_0");
    run_class<itl::gmres_solver<matrix_type>>("GMRES");
    run_class<itl::gmres_solver<matrix_type, ilu_0<matrix_type>>>("GMRES with ILU_0");
    run_class<itl::gmres_solver<matrix_type, ilu_0<matrix_type>, ilu_0<matrix_type>>>("GMRES with ILU_0 left and right");
    run_class<itl::qmr_solver<matrix_type>>("QMR");
    run_class<itl::qmr_solver<matrix_type, ilu_0<matrix_type>>>("QMR with ILU_0");
    run_class<itl::idr_s_solver<matrix_type>>("IDR(s)");
    run_class<itl::idr_s_solver<matrix_type, ilu_0<matrix_type>>>("IDR(s) with ILU_0");

    {   // gmres_test.cpp: one GMRES(1) step from x = 5 on the 2x2-grid Laplacian is exact
        matrix_type A(4, 4);
        laplacian_setup(A, 2, 2);
        mtl::dense_vector<double> b(4, 1.0), x(4, 5.0), r(4);
        itl::pc::identity<matrix_type> I(A);
        itl::basic_iteration<double> iter(b, 1, 1.e-8);
        gmres(A, x, b, I, I, iter, 1);
        r = A * x - b;
        MTL_THROW_IF(two_norm(r) > 0.00005, mtl::runtime_error("gmres doesn't converge"));
    }
    {   // qmr_test.cpp: 8x8-grid Laplacian, diagonal preconditioner from both sides
        const int N = 64;
        matrix_type A(N, N);
        laplacian_setup(A, 8, 8);
        itl::pc::diagonal<matrix_type> P(A);
        mtl::dense_vector<double> x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        itl::cyclic_iteration<double> iter(b, 500, 1.e-6, 0.0, 100);
        MTL_THROW_IF(qmr(A, x, b, P, P, iter) != 0, mtl::runtime_error("qmr"));
        mtl::dense_vector<double> r(b - A * x);
        MTL_THROW_IF(two_norm(r) > 1e-6 * two_norm(b), mtl::runtime_error("qmr residual"));
    }
    {   // idr_s_test.cpp: IDR(4) on the 10x10-grid Laplacian, counter reset first
        const int N = 100;
        matrix_type A(N, N);
        laplacian_setup(A, 10, 10);
        mtl::dense_vector<double> b(N), x(N, 1.0);
        b = A * x;
        itl::pc::identity<matrix_type> I(A);
        x = 0.0;
        mtl::seed<double>::counter = 0;
        itl::cyclic_iteration<double> iter(b, 3 * N, 1.e-8, 0.0, 50);
        MTL_THROW_IF(idr_s(A, x, b, I, I, iter, 4) != 0, mtl::runtime_error("idr_s"));
        MTL_THROW_IF(int(mtl::seed<double>::counter) != 4, mtl::runtime_error("seed counter"));
    }
    {   // ic_0_bicgstab_2_test.cpp
        const int N = 100;
        matrix_type A(N, N);
        laplacian_setup(A, 10, 10);
        itl::pc::ic_0<matrix_type> P(A);
        mtl::dense_vector<double> x(N, 1.0), b(N);
        b = A * x;
        x = 0;
        itl::cyclic_iteration<double> iter(b, 200, 1.e-6, 0.0, 1);
        MTL_THROW_IF(bicgstab_2(A, x, b, P, iter) != 0, mtl::runtime_error("bicgstab_2"));
    }
    return 0;
}
