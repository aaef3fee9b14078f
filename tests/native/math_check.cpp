// Host entry points over merzbild.jl_amd/csrc/mz_math.h for tests/test_math_cpu.py.
#include "../../merzbild.jl_amd/csrc/mz_math.h"

extern "C" void t_log(const double *x, double *y, long n) {
    for (long i = 0; i < n; ++i) y[i] = mz_log(x[i]);
}
extern "C" void t_pow(const double *g, double e, double *y, long n) {
    for (long i = 0; i < n; ++i) y[i] = mz_pow_pos(g[i], e);
}
extern "C" void t_sincos(const double *a, double *s, double *c, long n) {
    for (long i = 0; i < n; ++i) mz_sincos(a[i], s + i, c + i);
}
