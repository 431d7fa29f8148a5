// Host generators of the synthetic inputs (see mtl4gpu_gen.h).
#include "../../include/mtl4gpu_gen.h"

#include <cmath>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t index)
{
    return splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + index);
}

inline double unit53(uint64_t u) { return double(u >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)

inline double normal(uint64_t seed, uint64_t stream, uint64_t k)
{
    // Box-Muller on draws 2k and 2k+1
    double u1 = (double(draw(seed, stream, 2 * k) >> 11) + 1.0) * (1.0 / 9007199254740992.0);   // (0, 1]
    double u2 = unit53(draw(seed, stream, 2 * k + 1));
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

inline double harmonic(double a, double b) { return 2.0 * a * b / (a + b); }

struct CsrOut {
    int64_t* S;
    int64_t* I;
    double* V;
    int64_t pos = 0;
    void entry(int64_t c, double v)
    {
        if (I) I[pos] = c;
        if (V) V[pos] = v;
        pos++;
    }
    void end_row(int64_t r)
    {
        if (S) S[r + 1] = pos;
    }
};

}  // namespace

extern "C" {

void m4g_rng_uniform(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, double lo, double hi, double* out)
{
    for (int64_t k = 0; k < count; k++) out[k] = lo + (hi - lo) * unit53(draw(seed, stream, uint64_t(offset + k)));
}

void m4g_rng_normal(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, double* out)
{
    for (int64_t k = 0; k < count; k++) out[k] = normal(seed, stream, uint64_t(offset + k));
}

void m4g_rng_lognormal(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, double* out)
{
    for (int64_t k = 0; k < count; k++) out[k] = std::exp(normal(seed, stream, uint64_t(offset + k)));
}

void m4g_rng_int(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, int64_t lo, int64_t hi, int64_t* out)
{
    uint64_t span = uint64_t(hi - lo + 1);
    for (int64_t k = 0; k < count; k++) out[k] = lo + int64_t(draw(seed, stream, uint64_t(offset + k)) % span);
}

// laplacian_setup(A, m, n): row = i*n + j; 4 on the diagonal, -1 for the
// E/S/W/N neighbours; the inserter sorts each row by column.
int64_t m4g_gen_laplacian2d(int64_t m, int64_t n, int64_t* S, int64_t* I, double* V)
{
    CsrOut o{S, I, V};
    if (S) S[0] = 0;
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++) {
            int64_t row = i * n + j;
            if (i > 0) o.entry(row - n, -1.0);
            if (j > 0) o.entry(row - 1, -1.0);
            o.entry(row, 4.0);
            if (j < n - 1) o.entry(row + 1, -1.0);
            if (i < m - 1) o.entry(row + n, -1.0);
            o.end_row(row);
        }
    return o.pos;
}

int64_t m4g_gen_poisson3d(int64_t nx, int64_t ny, int64_t nz, int64_t* S, int64_t* I, double* V)
{
    CsrOut o{S, I, V};
    if (S) S[0] = 0;
    const int64_t sx = 1, sy = nx, sz = nx * ny;
    for (int64_t k = 0; k < nz; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                int64_t r = i + nx * (j + ny * k);
                if (k > 0) o.entry(r - sz, -1.0);
                if (j > 0) o.entry(r - sy, -1.0);
                if (i > 0) o.entry(r - sx, -1.0);
                o.entry(r, 6.0);
                if (i < nx - 1) o.entry(r + sx, -1.0);
                if (j < ny - 1) o.entry(r + sy, -1.0);
                if (k < nz - 1) o.entry(r + sz, -1.0);
                o.end_row(r);
            }
    return o.pos;
}

int64_t m4g_gen_convdiff27(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, double pe, int64_t* S, int64_t* I,
                           double* V)
{
    const int64_t n = nx * ny * nz;
    std::vector<double> kap;
    if (V) {
        kap.resize(n);
        m4g_rng_lognormal(seed, 2, 0, n, kap.data());
    }
    const double pex = pe * 1.0, pey = pe * 0.5, pez = pe * 0.25;
    CsrOut o{S, I, V};
    if (S) S[0] = 0;
    for (int64_t k = 0; k < nz; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                int64_t r = i + nx * (j + ny * k);
                double diag = 0.0;
                int64_t dpos = -1;
                for (int dk = -1; dk <= 1; dk++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int di = -1; di <= 1; di++) {
                            if (!di && !dj && !dk) {
                                dpos = o.pos;
                                o.entry(r, 0.0);
                                continue;
                            }
                            int taxi = std::abs(di) + std::abs(dj) + std::abs(dk);
                            double w = taxi == 1 ? 1.0 : (taxi == 2 ? 0.5 : 0.25);
                            int64_t ii = i + di, jj = j + dj, kk = k + dk;
                            bool inside = ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz;
                            int64_t c = ii + nx * (jj + ny * kk);
                            double kc = V ? (inside ? harmonic(kap[r], kap[c]) : kap[r]) : 0.0;
                            diag += w * kc;
                            if (!inside) continue;
                            double conv = 0.0;
                            if (taxi == 1) conv = di ? di * pex / 2 : (dj ? dj * pey / 2 : dk * pez / 2);
                            o.entry(c, -w * kc + conv * kc);
                        }
                if (V) V[dpos] = diag;
                o.end_row(r);
            }
    return o.pos;
}

int64_t m4g_gen_varcoef5(int64_t nx, int64_t ny, uint64_t seed, int64_t* S, int64_t* I, double* V)
{
    const int64_t n = nx * ny;
    std::vector<double> kap;
    if (V) {
        kap.resize(n);
        m4g_rng_lognormal(seed, 2, 0, n, kap.data());
    }
    CsrOut o{S, I, V};
    if (S) S[0] = 0;
    for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) {
            int64_t r = i + nx * j;
            double diag = 0.0;
            const int di[4] = {0, -1, 1, 0}, dj[4] = {-1, 0, 0, 1};   // sorted column order
            int64_t dpos = -1;
            for (int q = 0; q < 4; q++) {
                if (q == 2) {
                    dpos = o.pos;
                    o.entry(r, 0.0);
                }
                int64_t ii = i + di[q], jj = j + dj[q];
                bool inside = ii >= 0 && ii < nx && jj >= 0 && jj < ny;
                int64_t c = ii + nx * jj;
                double kc = V ? (inside ? harmonic(kap[r], kap[c]) : kap[r]) : 0.0;
                diag += kc;
                if (inside) o.entry(c, -kc);
            }
            if (V) V[dpos] = diag;
            o.end_row(r);
        }
    return o.pos;
}

int64_t m4g_gen_edge_laplacian(int64_t nx, int64_t ny, const double* theta, int64_t* n_edges, int64_t* S,
                               int64_t* I, double* V, int64_t* ei, int64_t* ej, int64_t* aptr, int64_t* aidx,
                               double* acoef)
{
    // edges: horizontal (i,j)-(i+1,j), then vertical (i,j)-(i,j+1), then
    // boundary edges (one per missing neighbour, west/east/south/north order)
    const int64_t nh = (nx - 1) * ny, nv = nx * (ny - 1);
    auto hedge = [&](int64_t i, int64_t j) { return j * (nx - 1) + i; };
    auto vedge = [&](int64_t i, int64_t j) { return nh + j * nx + i; };
    std::vector<int64_t> bstart(nx * ny + 1, 0);
    for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) {
            int64_t r = i + nx * j;
            bstart[r + 1] = bstart[r] + (i == 0) + (i == nx - 1) + (j == 0) + (j == ny - 1);
        }
    const int64_t nb = bstart[nx * ny];
    if (n_edges) *n_edges = nh + nv + nb;
    if (ei || ej) {
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i + 1 < nx; i++) {
                if (ei) ei[hedge(i, j)] = i + nx * j;
                if (ej) ej[hedge(i, j)] = i + 1 + nx * j;
            }
        for (int64_t j = 0; j + 1 < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                if (ei) ei[vedge(i, j)] = i + nx * j;
                if (ej) ej[vedge(i, j)] = i + nx * (j + 1);
            }
        for (int64_t r = 0; r < nx * ny; r++)
            for (int64_t b = bstart[r]; b < bstart[r + 1]; b++) {
                if (ei) ei[nh + nv + b] = r;
                if (ej) ej[nh + nv + b] = -1;
            }
    }
    CsrOut o{S, I, V};
    if (S) S[0] = 0;
    int64_t apos = 0;
    if (aptr) aptr[0] = 0;
    auto add = [&](int64_t e, double c, double& acc) {
        if (theta) acc += c * theta[e];
        if (aidx) aidx[apos] = e;
        if (acoef) acoef[apos] = c;
        apos++;
    };
    for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) {
            int64_t r = i + nx * j;
            const int di[4] = {0, -1, 1, 0}, dj[4] = {-1, 0, 0, 1};
            for (int q = 0; q < 5; q++) {
                double acc = 0.0;
                if (q == 2) {   // diagonal: every incident edge with +1
                    if (i > 0) add(hedge(i - 1, j), 1.0, acc);
                    if (i + 1 < nx) add(hedge(i, j), 1.0, acc);
                    if (j > 0) add(vedge(i, j - 1), 1.0, acc);
                    if (j + 1 < ny) add(vedge(i, j), 1.0, acc);
                    for (int64_t b = bstart[r]; b < bstart[r + 1]; b++) add(nh + nv + b, 1.0, acc);
                    o.entry(r, acc);
                } else {
                    int qq = q < 2 ? q : q - 1;
                    int64_t ii = i + di[qq], jj = j + dj[qq];
                    if (ii < 0 || ii >= nx || jj < 0 || jj >= ny) continue;
                    int64_t e = di[qq] ? hedge(std::min(i, ii), j) : vedge(i, std::min(j, jj));
                    add(e, -1.0, acc);
                    o.entry(ii + nx * jj, acc);
                }
                if (aptr) aptr[o.pos] = apos;
            }
            o.end_row(r);
        }
    return o.pos;
}

}  // extern "C"
