// Expectation counter shared by the drop-in programs: every failed check is
// printed, and the count becomes the process exit status (0: all passed).
#pragma once
#include <cmath>
#include <cstdio>

struct Checks {
    int failed = 0;
    void expect(bool ok, const char* what)
    {
        if (!ok) {
            ++failed;
            std::printf("FAILED: %s\n", what);
        }
    }
    void near(double got, double want, double tol, const char* what)
    {
        if (!(std::fabs(got - want) <= tol)) {
            ++failed;
            std::printf("FAILED: %s: %.17g, expected %.17g (tolerance %.1e)\n", what, got, want, tol);
        }
    }
    int status() const
    {
        if (failed) std::printf("%d check(s) failed\n", failed);
        else std::printf("all checks passed\n");
        return failed ? 1 : 0;
    }
};
