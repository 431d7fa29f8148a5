// Host-side compilation of triangular sweeps into tile-pipelined schedules.
//
// A sweep (forward on L or backward on U) is split into tiles of rows.  Each
// tile is executed by one workgroup that keeps the tile's vector entries in
// LDS and walks the tile's local dependency levels in groups of at most 64
// rows, one wavefront pass per group.
// Values needed by later tiles are exported through a device exchange buffer.
// Tiles are numbered in a topological order of the tile graph and handed to
// workgroups by an atomic ticket, so every wait of a tile is on a tile with a
// smaller ticket, which already runs or has finished: the pipeline cannot
// deadlock whatever the residency.
//
// For matrices on a natural-order grid (inferred from the pattern) the tiles
// are boxes in skewed coordinates t1 = k, t2 = j + a*k, t3 = i + b*j + c*k in
// which every dependency is monotone; otherwise tiles are blocks of
// consecutive rows.  Neither choice changes any arithmetic: the operation
// order per row is the reference's (lower_trisolve.hpp:191-209 / 172-189,
// upper_trisolve.hpp:159-176).
#include <sys/mman.h>

#include "schedule.hpp"

#include <thread>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <numeric>
#include <set>
#include <stdexcept>

namespace m4g {

void* large_block_alloc(size_t bytes)
{
    constexpr size_t huge = size_t(2) << 20;
    const size_t rounded = (bytes + huge - 1) / huge * huge;
    void* p = std::aligned_alloc(huge, rounded);
    if (!p) throw std::bad_alloc();
    madvise(p, rounded, MADV_HUGEPAGE);   // advice only: ignored where huge pages are off
    return p;
}

void large_block_free(void* p) noexcept { std::free(p); }

// Host threads a sweep plan may use: M4G_PLAN_THREADS, else half the
// hardware threads (the two directions of a preconditioner are planned at
// once), at most 16.
int plan_threads()
{
    static const int nt = [] {
        if (const char* e = std::getenv("M4G_PLAN_THREADS"); e && std::atoi(e) > 0)
            return std::min(64, std::atoi(e));
        const int hw = int(std::thread::hardware_concurrency());
        return std::max(1, std::min(16, hw / 2));
    }();
    return nt;
}

// Runs fn(begin, end, thread) over [0, count) in contiguous chunks of at
// least min_chunk items on up to nt host threads (the caller takes the
// first chunk).  fn must not throw.
template <typename F>
static void parallel_chunks(int64_t count, int nt, F&& fn, int64_t min_chunk = 16384)
{
    const int64_t k = std::max<int64_t>(1, std::min<int64_t>(nt, count / std::max<int64_t>(1, min_chunk)));
    if (k <= 1) {
        if (count > 0) fn(int64_t(0), count, 0);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(size_t(k - 1));
    for (int64_t t = 1; t < k; t++)
        th.emplace_back([&fn, count, k, t] { fn(count * t / k, count * (t + 1) / k, int(t)); });
    fn(int64_t(0), count / k, 0);
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// grid inference
// ---------------------------------------------------------------------------
// Bit set of the neighbour offsets (di, dj, dk) that rows r_begin, r_begin +
// stride, ... < r_end use on the grid, or -1 when some entry couples cells
// further apart than one step.
static int64_t grid_score_rows(int64_t n, const int64_t* S, const int64_t* I, int64_t nx, int64_t ny, int64_t stride,
                               int64_t r_begin, int64_t r_end)
{
    const int64_t nxy = nx * ny, nz = n / nxy;
    uint32_t seen = 0;   // bit (di+1) + 3(dj+1) + 9(dk+1)
    // the 27 neighbour offsets; c - r = off with the neighbour inside the
    // grid identifies (di, dj, dk) uniquely (mixed-radix digits)
    int64_t off[27];
    for (int code = 0; code < 27; code++)
        off[code] = (code % 3 - 1) + nx * ((code / 3) % 3 - 1) + nxy * (code / 9 - 1);
    // offset -> code table (open addressing); an offset that two codes share
    // (nx or ny below 3) is marked ambiguous and searched code by code
    struct Slot {
        int64_t off;
        int code;   // -1 empty, -2 ambiguous
    };
    Slot tab[64];
    for (Slot& sl : tab) sl = Slot{0, -1};
    auto slot_of = [&](int64_t d) -> Slot& {
        size_t h = size_t((uint64_t(d) * 0x9E3779B97F4A7C15ull) >> 58);
        while (tab[h].code != -1 && tab[h].off != d) h = (h + 1) & 63;
        return tab[h];
    };
    for (int code = 0; code < 27; code++) {
        Slot& sl = slot_of(off[code]);
        if (sl.code == -1) sl = Slot{off[code], code};
        else sl.code = -2;
    }
    int64_t ri = r_begin % nx, rj = (r_begin / nx) % ny, rk = r_begin / nxy;
    for (int64_t r = r_begin; r < r_end; r += stride) {
        if (stride != 1) {
            ri = r % nx;
            rj = (r / nx) % ny;
            rk = r / nxy;
        }
        int guess = 0;
        for (int64_t p = S[r]; p < S[r + 1]; p++) {
            const int64_t d = I[p] - r;
            const Slot& sl = slot_of(d);
            if (sl.code == -1) return -1;   // no neighbour has this offset
            if (sl.code >= 0) {
                const int code = sl.code;
                const int64_t ci = ri + code % 3 - 1, cj = rj + (code / 3) % 3 - 1, ck = rk + code / 9 - 1;
                if (!(ci >= 0 && ci < nx && cj >= 0 && cj < ny && ck >= 0 && ck < nz)) return -1;
                seen |= uint32_t(1) << code;
                guess = code + 1;
                continue;
            }
            int found = -1;
            for (int t = 0; t < 27 && found < 0; t++) {
                const int code = (guess + t) % 27;
                if (off[code] != d) continue;
                const int64_t ci = ri + code % 3 - 1, cj = rj + (code / 3) % 3 - 1, ck = rk + code / 9 - 1;
                if (ci >= 0 && ci < nx && cj >= 0 && cj < ny && ck >= 0 && ck < nz) found = code;
            }
            if (found < 0) return -1;
            seen |= uint32_t(1) << found;
            guess = found + 1;   // columns ascend: the next entry's offset is larger
        }
        if (stride == 1 && ++ri == nx) {   // next row's coordinates, stepped
            ri = 0;
            if (++rj == ny) {
                rj = 0;
                ++rk;
            }
        }
    }
    return int64_t(seen);
}

// Number of distinct neighbour offsets (di, dj, dk) of the pattern on the
// grid, or -1 when some entry couples cells further apart than one step;
// every stride-th row is looked at (rows split over the planning threads).
static int64_t grid_score(int64_t n, const int64_t* S, const int64_t* I, int64_t nx, int64_t ny, int64_t stride)
{
    const int64_t rows = (n + stride - 1) / stride;
    const int nt = rows >= 65536 ? plan_threads() : 1;
    std::vector<int64_t> part(size_t(nt), 0);
    parallel_chunks(rows, nt, [&](int64_t a, int64_t b, int th) {
        part[size_t(th)] = grid_score_rows(n, S, I, nx, ny, stride, a * stride, std::min(n, b * stride));
    });
    uint32_t seen = 0;
    for (int64_t v : part) {
        if (v < 0) return -1;
        seen |= uint32_t(v);
    }
    return int64_t(__builtin_popcount(seen));
}

GridInfo infer_grid(int64_t n, const int64_t* S, const int64_t* I)
{
    GridInfo best;
    if (n < 4) return best;
    std::set<int64_t> offs;
    {
        // distinct lower offsets r - c, per chunk of rows a short list scanned
        // linearly; more than 64 means no stencil pattern
        const int nt = n >= 65536 ? plan_threads() : 1;
        std::vector<std::vector<int64_t>> seen{size_t(nt)};
        parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int th) {
            std::vector<int64_t>& sv = seen[size_t(th)];
            int64_t recent[64];   // direct-mapped memo of offsets already listed
            std::fill(recent, recent + 64, int64_t(0));
            for (int64_t r = r0; r < r1 && sv.size() <= 64; r++)
                for (int64_t p = S[r]; p < S[r + 1]; p++)
                    if (I[p] < r) {
                        const int64_t d = r - I[p];
                        int64_t& m = recent[(uint64_t(d) * 0x9E3779B97F4A7C15ull) >> 58];
                        if (m == d) continue;
                        m = d;
                        if (std::find(sv.begin(), sv.end(), d) == sv.end()) sv.push_back(d);
                    }
        });
        for (const auto& sv : seen) {
            offs.insert(sv.begin(), sv.end());
            if (offs.size() > 64) return best;
        }
    }
    if (offs.empty()) return best;
    std::set<int64_t> nxs;
    for (int64_t d : offs)
        for (int64_t e = -1; e <= 1; e++) {
            int64_t nx = d + e;
            if (nx >= 2 && nx <= n && n % nx == 0) nxs.insert(nx);
        }
    // every candidate scored on a sample of rows; then the full check, best
    // candidate first: fewest distinct offsets (the tightest stencil), then
    // the larger nx, then the later candidate (the larger nx*ny)
    const int64_t sample = std::max<int64_t>(1, n / 4096);
    struct Cand {
        int64_t score, nx, ny, nz, order;
    };
    std::vector<Cand> cands;
    for (int64_t nx : nxs) {
        std::set<int64_t> nxys;
        for (int64_t d : offs) {
            if (d < nx) continue;
            for (int64_t e = -nx - 1; e <= nx + 1; e++) {
                int64_t v = d + e;
                if (v > nx && v < n && v % nx == 0 && n % v == 0) nxys.insert(v);
            }
        }
        nxys.insert(n);   // 2D (or 1D when nx == n)
        for (int64_t nxy : nxys) {
            const int64_t ny = nxy / nx, nz = n / nxy;
            const int64_t sc = grid_score(n, S, I, nx, ny, sample);
            if (sc >= 0) cands.push_back({sc, nx, ny, nz, int64_t(cands.size())});
        }
    }
    std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) {
        if (a.score != b.score) return a.score < b.score;
        if (a.nx != b.nx) return a.nx > b.nx;
        return a.order > b.order;
    });
    for (const Cand& c : cands)
        if (grid_score(n, S, I, c.nx, c.ny, 1) >= 0) {
            best.nx = c.nx;
            best.ny = c.ny;
            best.nz = c.nz;
            break;
        }
    return best;
}

// ---------------------------------------------------------------------------
// levels
// ---------------------------------------------------------------------------
std::vector<int32_t> natural_levels(const SweepRows& s, int64_t* nlevels)
{
    std::vector<int32_t> lvl(s.n, 0);
    int32_t maxl = -1;
    for (int64_t q = 0; q < s.n; q++) {
        int64_t r = s.lower ? q : s.n - 1 - q;
        int32_t l = 0;
        for (int64_t p = s.ptr[r]; p < s.ptr[r + 1]; p++) l = std::max(l, lvl[s.col[p]] + 1);
        lvl[r] = l;
        maxl = std::max(maxl, l);
    }
    if (nlevels) *nlevels = maxl + 1;
    return lvl;
}

// ---------------------------------------------------------------------------
// tile coordinates
// ---------------------------------------------------------------------------
struct Coef {
    bool ok = false;
    int64_t a = 0, b = 0, c = 0;
};

static Coef skew_coefficients(const SweepRows& s, const GridInfo& g)
{
    Coef k;
    if (!g.nx) return k;
    const int64_t nx = g.nx, ny = g.ny, nxy = g.nx * g.ny;
    int64_t a = 0, b = 0;
    // distinct (di, dj, dk) of the dependencies, collected per chunk of rows
    const int nt = s.n >= 65536 ? plan_threads() : 1;
    std::vector<std::vector<std::array<int, 3>>> part{size_t(nt)};
    parallel_chunks(s.n, nt, [&](int64_t r0, int64_t r1, int th) {
        std::vector<std::array<int, 3>>& dl = part[size_t(th)];
        // the offset c - r of an entry repeats from row to row: a small table
        // maps it to the (di, dj, dk) last seen with it, which is that entry's
        // as soon as it lands inside the grid from r (digits in range make
        // the mixed-radix form c - r = di + nx*dj + nxy*dk unique)
        struct Slot {
            int64_t off;
            int d[3];
            bool used;
        };
        Slot tab[64] = {};
        int64_t ri = r0 % nx, rj = (r0 / nx) % ny, rk = r0 / nxy;
        for (int64_t r = r0; r < r1; r++) {
            for (int64_t p = s.ptr[r]; p < s.ptr[r + 1]; p++) {
                const int64_t c = s.col[p], off = c - r;
                Slot& sl = tab[(uint64_t(off) * 0x9E3779B97F4A7C15ull) >> 58];
                if (sl.used && sl.off == off) {
                    const int64_t ci = ri + sl.d[0], cj = rj + sl.d[1], ck = rk + sl.d[2];
                    if (ci >= 0 && ci < nx && cj >= 0 && cj < ny && ck >= 0 && ck < g.nz) continue;
                }
                int di = int(c % nx - ri), dj = int((c / nx) % ny - rj), dk = int(c / nxy - rk);
                sl = Slot{off, {di, dj, dk}, true};
                if (!s.lower) { di = -di; dj = -dj; dk = -dk; }
                const std::array<int, 3> d = {di, dj, dk};
                if (std::find(dl.begin(), dl.end(), d) == dl.end()) dl.push_back(d);
            }
            if (++ri == nx) {
                ri = 0;
                if (++rj == ny) {
                    rj = 0;
                    ++rk;
                }
            }
        }
    });
    std::vector<std::array<int, 3>> deltas;
    for (const auto& dl : part)
        for (const auto& d : dl)
            if (std::find(deltas.begin(), deltas.end(), d) == deltas.end()) deltas.push_back(d);
    for (auto& d : deltas) {
        if (d[2] > 0) return k;
        if (d[2] == 0 && d[1] > 0) return k;
        if (d[2] == 0 && d[1] == 0 && d[0] >= 0) return k;
        if (d[2] < 0) a = std::max<int64_t>(a, d[1]);
        if (d[2] == 0 && d[1] < 0) b = std::max<int64_t>(b, d[0]);
    }
    int64_t c = 0;
    for (auto& d : deltas)
        if (d[2] < 0) c = std::max<int64_t>(c, d[0] + b * d[1]);
    k.ok = true;
    k.a = a;
    k.b = b;
    k.c = c;
    return k;
}

struct TileCoords {
    int64_t n1 = 1, n2 = 1, n3 = 1;   // tile counts per coordinate
};

static inline void grid_coords(const GridInfo& g, int64_t r, int64_t& i, int64_t& j, int64_t& k)
{
    i = r % g.nx;
    j = (r / g.nx) % g.ny;
    k = r / (g.nx * g.ny);
}

// Wavefront passes per level of a box tile of extents x1, x2, x3 in which
// cell (t1, t2, t3) sits on local level w1*t1 + w2*t2 + w3*t3: a level of
// more than 64 cells takes several passes.  Returns passes / levels.
static double passes_per_level(int64_t x1, int64_t x2, int64_t x3, int64_t w1, int64_t w2, int64_t w3)
{
    std::vector<int64_t> cnt(size_t(w1 * (x1 - 1) + w2 * (x2 - 1) + w3 * (x3 - 1) + 1), 0);
    for (int64_t t1 = 0; t1 < x1; t1++)
        for (int64_t t2 = 0; t2 < x2; t2++) {
            const int64_t base = w1 * t1 + w2 * t2;
            for (int64_t t3 = 0; t3 < x3; t3++) cnt[size_t(base + w3 * t3)]++;
        }
    int64_t groups = 0, levels = 0;
    for (int64_t c : cnt)
        if (c > 0) {
            groups += (c + 63) / 64;
            levels++;
        }
    return levels ? double(groups) / double(levels) : 1.0;
}

TileShape choose_tiles(const SweepRows& s, const GridInfo& g, int64_t cap, double hop_us, double step_us,
                       const std::vector<int32_t>* levels)
{
    TileShape best;
    Coef cf = skew_coefficients(s, g);
    if (!cf.ok) {
        best.grid = false;
        best.rows_per_block = std::min<int64_t>(cap, std::max<int64_t>(s.n, 1));
        return best;
    }
    // level function of the natural DAG is linear on a grid: measure it
    int64_t nlev = 0;
    std::vector<int32_t> own;
    if (!levels) own = natural_levels(s, &nlev);
    const std::vector<int32_t>& lvl = levels ? *levels : own;
    if (levels)
        for (int32_t l : lvl) nlev = std::max<int64_t>(nlev, int64_t(l) + 1);
    auto at = [&](int64_t i, int64_t j, int64_t k) -> double {
        int64_t r = i + g.nx * (j + g.ny * k);
        return r < s.n ? double(lvl[r]) : 0.0;
    };
    double base = at(0, 0, 0);
    double al = g.nx > 1 ? std::fabs(at(1, 0, 0) - base) : 0;
    double be = g.ny > 1 ? std::fabs(at(0, 1, 0) - base) : 0;
    double ga = g.nz > 1 ? std::fabs(at(0, 0, 1) - base) : 0;
    if (!s.lower) {   // levels count from the last row
        al = g.nx > 1 ? std::fabs(at(g.nx - 2, g.ny - 1, g.nz - 1) - at(g.nx - 1, g.ny - 1, g.nz - 1)) : 0;
        be = g.ny > 1 ? std::fabs(at(g.nx - 1, g.ny - 2, g.nz - 1) - at(g.nx - 1, g.ny - 1, g.nz - 1)) : 0;
        ga = g.nz > 1 ? std::fabs(at(g.nx - 1, g.ny - 1, g.nz - 2) - at(g.nx - 1, g.ny - 1, g.nz - 1)) : 0;
    }
    // level = al*t3 + (be - al*b)*t2 + (ga - al*c - a*(be - al*b))*t1
    double w3 = al, w2 = std::fabs(be - al * cf.b), w1 = std::fabs(ga - al * cf.c - cf.a * (be - al * cf.b));
    int64_t ext1 = g.nz, ext2 = (g.ny - 1) + cf.a * (g.nz - 1) + 1,
            ext3 = (g.nx - 1) + cf.b * (g.ny - 1) + cf.c * (g.nz - 1) + 1;
    std::vector<int64_t> c1, c2, c3;
    for (int64_t e = 1; e < ext1; e *= 2) c1.push_back(e);
    c1.push_back(ext1);
    for (int64_t e = 1; e < ext2; e *= 2) c2.push_back(e);
    c2.push_back(ext2);
    for (int64_t e = 16; e < ext3; e *= 2) c3.push_back(e);
    c3.push_back(ext3);
    double best_cost = std::numeric_limits<double>::infinity();
    for (int64_t e1 : c1)
        for (int64_t e2 : c2)
            for (int64_t e3 : c3) {
                // rows in a tile: at most e1 planes x e2 lines x min(e3, nx) cells
                int64_t rows = e1 * std::min<int64_t>(e2, g.ny) * std::min<int64_t>(e3, g.nx);
                if (rows > cap) continue;
                int64_t n1 = (ext1 + e1 - 1) / e1, n2 = (ext2 + e2 - 1) / e2, n3 = (ext3 + e3 - 1) / e3;
                double hops = double(n1 - 1 + n2 - 1 + n3 - 1);
                // every global level lies on the critical path whatever the
                // tiling, at the pace of the tile's wavefront passes per level;
                // tiles add a hop latency at each boundary crossed
                double passes = passes_per_level(std::min(e1, ext1), std::min(e2, ext2), std::min(e3, ext3),
                                                 int64_t(std::llround(w1)), int64_t(std::llround(w2)),
                                                 int64_t(std::llround(w3)));
                double cost = hops * hop_us + double(nlev) * passes * step_us;
                if (cost < best_cost) {
                    best_cost = cost;
                    best.grid = true;
                    best.a = cf.a;
                    best.b = cf.b;
                    best.c = cf.c;
                    best.e1 = e1;
                    best.e2 = e2;
                    best.e3 = e3;
                }
            }
    if (!best.grid) {
        best.rows_per_block = std::min<int64_t>(cap, std::max<int64_t>(s.n, 1));
    }
    return best;
}

// ---------------------------------------------------------------------------
// plan construction
// ---------------------------------------------------------------------------
int sweep_piece_width(int64_t max_len)
{
    for (int w : {3, 4, 8})
        if (max_len <= w) return w;
    return 16;
}

SweepPlan build_sweep_plan(const SweepRows& s, const GridInfo& g, const TileShape& shape, int64_t cap, int width,
                           int64_t levels)
{
    SweepPlan p;
    const int64_t n = s.n;
    p.n = n;
    if (width <= 0 && std::getenv("M4G_SWEEP_WIDTH")) width = std::atoi(std::getenv("M4G_SWEEP_WIDTH"));
    if (width <= 0) {
        int64_t max_len = 0;
        for (int64_t r = 0; r < n; r++) max_len = std::max(max_len, s.ptr[r + 1] - s.ptr[r]);
        width = sweep_piece_width(max_len);
    }
    p.width = width;
    p.lower = s.lower;
    p.unit = true;
    for (int64_t r = 0; r < n; r++)
        if (s.diag_pos[r] >= 0) { p.unit = false; break; }
    p.shape = shape;
    if (levels >= 0) p.levels = levels;
    else natural_levels(s, &p.levels);
    const int nt = plan_threads();

    // 1. tile of every row: linear box index, then tickets = the occupied
    //    boxes in topological order (ascending box index for the lower
    //    sweep, descending for the upper)
    PlanVec32 tile(n);
    int64_t nbox = 1;
    PlanVec64 lin(n);
    if (shape.grid) {
        const int64_t N1 = (g.nz - 1) / shape.e1 + 1;
        const int64_t N2 = ((g.ny - 1) + shape.a * (g.nz - 1)) / shape.e2 + 1;
        const int64_t N3 = ((g.nx - 1) + shape.b * (g.ny - 1) + shape.c * (g.nz - 1)) / shape.e3 + 1;
        nbox = N1 * N2 * N3;
        const int64_t nxy = g.nx * g.ny;
        // along a grid line (fixed j, k) t3 grows by one per row: its box
        // index is stepped, not divided
        parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int) {
            for (int64_t r = r0; r < r1;) {
                const int64_t k = r / nxy, j = (r - k * nxy) / g.nx, i = r - k * nxy - j * g.nx;
                const int64_t t1 = k, t2 = j + shape.a * k, t3 = i + shape.b * j + shape.c * k;
                const int64_t head = ((t1 / shape.e1) * N2 + t2 / shape.e2) * N3;
                int64_t q3 = t3 / shape.e3, m3 = t3 - q3 * shape.e3;
                const int64_t line_end = std::min(r1, r + (g.nx - i));
                for (; r < line_end; r++) {
                    lin[r] = head + q3;
                    if (++m3 == shape.e3) {
                        m3 = 0;
                        ++q3;
                    }
                }
            }
        });
    } else {
        const int64_t rb = std::max<int64_t>(1, shape.rows_per_block);
        nbox = n ? (n - 1) / rb + 1 : 1;
        parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int) {
            for (int64_t r = r0; r < r1; r++) lin[r] = r / rb;
        });
    }
    {
        std::vector<int32_t> rank(size_t(nbox), -1);
        parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int) {
            for (int64_t r = r0; r < r1; r++) __atomic_store_n(&rank[size_t(lin[r])], 0, __ATOMIC_RELAXED);
        });
        p.ntiles = 0;
        for (int64_t b = 0; b < nbox; b++) {
            int32_t& rk = rank[size_t(s.lower ? b : nbox - 1 - b)];
            if (rk == 0) rk = int32_t(p.ntiles++);
        }
        parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int) {
            for (int64_t r = r0; r < r1; r++) tile[r] = rank[size_t(lin[r])];
        });
    }
    PlanVec64().swap(lin);
    const int64_t T = p.ntiles;

    // 2. rows of each tile in natural order: per-chunk counts per tile, so
    //    every chunk writes its rows at fixed positions
    p.tile_row_begin.assign(T + 1, 0);
    p.rows.resize(n);
    {
        const int64_t nc = std::max<int64_t>(1, std::min<int64_t>(nt, n / 4096));
        std::vector<std::vector<int32_t>> cnt(size_t(nc), std::vector<int32_t>(size_t(T), 0));
        auto chunk = [&](int64_t c, int64_t& r0, int64_t& r1) { r0 = n * c / nc; r1 = n * (c + 1) / nc; };
        parallel_chunks(nc, nt, [&](int64_t c0, int64_t c1, int) {
            for (int64_t c = c0; c < c1; c++) {
                int64_t r0, r1;
                chunk(c, r0, r1);
                for (int64_t r = r0; r < r1; r++) cnt[c][tile[r]]++;
            }
        }, 1);
        for (int64_t t = 0; t < T; t++) {
            int32_t tot = 0;
            for (int64_t c = 0; c < nc; c++) tot += cnt[c][t];
            p.tile_row_begin[t + 1] = p.tile_row_begin[t] + tot;
        }
        // cnt becomes each chunk's first position per tile
        for (int64_t t = 0; t < T; t++) {
            int32_t pos = p.tile_row_begin[t];
            for (int64_t c = 0; c < nc; c++) {
                const int32_t k = cnt[c][t];
                cnt[c][t] = pos;
                pos += k;
            }
        }
        parallel_chunks(nc, nt, [&](int64_t c0, int64_t c1, int) {
            for (int64_t c = c0; c < c1; c++) {
                int64_t r0, r1;
                chunk(c, r0, r1);
                for (int64_t r = r0; r < r1; r++) p.rows[cnt[c][tile[r]]++] = int32_t(r);
            }
        }, 1);
    }
    for (int64_t t = 0; t < T; t++)
        p.max_tile_rows = std::max<int64_t>(p.max_tile_rows, p.tile_row_begin[t + 1] - p.tile_row_begin[t]);
    if (p.max_tile_rows > cap) throw std::runtime_error("sweep tile exceeds LDS capacity");

    // 3. per tile: pieces in processing order (ascending rows for the lower
    //    sweep, descending for the upper) with their local levels, then
    //    ordered stably by level and cut into groups of at most 64 pieces of
    //    one level.  Rows read by other tiles are marked exported.
    const int W = p.width;
    struct Piece {
        int32_t row, k0, len, level;
    };
    std::vector<std::vector<Piece>> tpieces{size_t(T)};
    std::vector<std::vector<int32_t>> tgroups{size_t(T)};   // tile-local first piece of each group
    PlanVec32 llev(n);   // local level of the row's last piece (written before it is read)
    std::vector<char> exported(n, 0);
    parallel_chunks(T, nt, [&](int64_t ta, int64_t tb, int) {
        std::vector<Piece> raw;
        std::vector<int32_t> lc;
        for (int64_t t = ta; t < tb; t++) {
            const int32_t q0 = p.tile_row_begin[t], q1 = p.tile_row_begin[t + 1];
            raw.clear();
            int32_t maxlev = 0;
            for (int32_t qq = 0; qq < q1 - q0; qq++) {
                const int32_t r = p.rows[s.lower ? q0 + qq : q1 - 1 - qq];
                const int64_t len = s.ptr[r + 1] - s.ptr[r];
                int32_t prev = -1;
                for (int64_t k0 = 0; k0 == 0 || k0 < len; k0 += W) {
                    const int64_t k1 = std::min<int64_t>(len, k0 + W);
                    int32_t l = prev + 1;
                    for (int64_t e = s.ptr[r] + k0; e < s.ptr[r] + k1; e++) {
                        const int64_t c = s.col[e];
                        if (tile[c] == t) l = std::max(l, llev[c] + 1);
                        else __atomic_store_n(&exported[c], char(1), __ATOMIC_RELAXED);
                    }
                    raw.push_back({r, int32_t(k0), int32_t(k1 - k0), l});
                    prev = l;
                }
                llev[r] = prev;
                maxlev = std::max(maxlev, prev);
            }
            // stable counting sort by level
            lc.assign(size_t(maxlev) + 2, 0);
            for (const Piece& pc : raw) lc[size_t(pc.level) + 1]++;
            for (size_t l = 1; l < lc.size(); l++) lc[l] += lc[l - 1];
            std::vector<Piece>& out = tpieces[size_t(t)];
            out.resize(raw.size());
            for (const Piece& pc : raw) out[size_t(lc[size_t(pc.level)]++)] = pc;
            std::vector<int32_t>& grp = tgroups[size_t(t)];
            grp.clear();
            for (int32_t q = 0; q < int32_t(out.size());) {
                int32_t q1e = q;
                while (q1e < int32_t(out.size()) && q1e - q < 64 && out[q1e].level == out[q].level) q1e++;
                grp.push_back(q);
                q = q1e;
            }
        }
    }, 1);
    // global numbering of pieces and groups: tiles in ticket order
    std::vector<int64_t> tpiece0(size_t(T) + 1, 0), tgroup0(size_t(T) + 1, 0);
    for (int64_t t = 0; t < T; t++) {
        tpiece0[t + 1] = tpiece0[t] + int64_t(tpieces[t].size());
        tgroup0[t + 1] = tgroup0[t] + int64_t(tgroups[t].size());
    }
    p.npieces = tpiece0[T];
    p.ngroups = tgroup0[T];
    if (p.npieces >= (int64_t(1) << 31) / std::max(W, 1)) throw std::runtime_error("sweep too large");

    p.group_piece_begin.assign(size_t(p.ngroups) + 1, int32_t(p.npieces));
    p.tile_group_begin.assign(T + 1, int32_t(p.ngroups));
    for (int64_t t = 0; t < T; t++) {
        p.tile_group_begin[t] = int32_t(tgroup0[t]);
        p.max_steps = std::max<int64_t>(p.max_steps, tgroup0[t + 1] - tgroup0[t]);
    }

    // 4. per tile: LDS slots in (group, lane) order (the rows of a group, and
    //    on a grid the rows they read from the group before, sit in
    //    consecutive slots, so a wavefront's LDS accesses fall in distinct
    //    banks); the tile's export count, in the order its rows finish
    std::vector<int32_t> slot(n, -1);
    std::vector<int32_t> texports(size_t(T), 0);
    parallel_chunks(T, nt, [&](int64_t ta, int64_t tb, int) {
        for (int64_t t = ta; t < tb; t++) {
            int32_t next = 0, nx = 0;
            for (const Piece& pc : tpieces[t]) {
                if (slot[pc.row] < 0) slot[pc.row] = next++;
                if (exported[pc.row] && pc.k0 + pc.len == s.ptr[pc.row + 1] - s.ptr[pc.row]) nx++;
            }
            texports[t] = nx;
        }
    }, 1);
    p.row_slot.resize(n);
    parallel_chunks(n, nt, [&](int64_t q0, int64_t q1, int) {
        for (int64_t q = q0; q < q1; q++) p.row_slot[q] = slot[p.rows[q]];
    });
    // export slots: contiguous per tile, in the order the rows finish (group
    // order of their last pieces); each tile's list of exported LDS slots, and
    // per group the tile's exports finished before it
    p.tile_export_begin.assign(T + 1, 0);
    for (int64_t t = 0; t < T; t++) {
        p.tile_export_begin[t + 1] = p.tile_export_begin[t] + texports[t];
        p.max_tile_exports = std::max<int64_t>(p.max_tile_exports, texports[t]);
    }
    p.nexport = p.tile_export_begin[T];
    p.xlist.assign(size_t(p.nexport), 0);
    p.group_xbefore.assign(size_t(p.ngroups), 0);
    std::vector<int32_t> xslot(n, -1);
    parallel_chunks(T, nt, [&](int64_t ta, int64_t tb, int) {
        for (int64_t t = ta; t < tb; t++) {
            const int32_t xb = p.tile_export_begin[t];
            int32_t x = xb;
            const std::vector<Piece>& tp = tpieces[t];
            const std::vector<int32_t>& grp = tgroups[t];
            for (size_t gl = 0; gl < grp.size(); gl++) {
                const int64_t gi = tgroup0[t] + int64_t(gl);
                p.group_piece_begin[gi] = int32_t(tpiece0[t] + grp[gl]);
                p.group_xbefore[gi] = x - xb;
                const size_t qe = gl + 1 < grp.size() ? size_t(grp[gl + 1]) : tp.size();
                for (size_t q = size_t(grp[gl]); q < qe; q++) {
                    const Piece& pc = tp[q];
                    if (exported[pc.row] && pc.k0 + pc.len == s.ptr[pc.row + 1] - s.ptr[pc.row]) {
                        xslot[pc.row] = x;
                        p.xlist[x++] = slot[pc.row];
                    }
                }
            }
        }
    }, 1);
    // tiles without rows: empty group range
    for (int64_t t = T - 1; t >= 0; t--)
        if (p.tile_row_begin[t + 1] == p.tile_row_begin[t]) p.tile_group_begin[t] = p.tile_group_begin[t + 1];

    // 5. piece records in group order; sources still as slot / -(export + 1)
    p.pc_slot.resize(p.npieces);
    p.pc_len.resize(p.npieces);
    p.pc_export.resize(p.npieces);
    p.pc_diag_pos.resize(p.npieces);
    p.pc_src.resize(size_t(p.npieces) * W);
    p.pc_val_pos.resize(size_t(p.npieces) * W);
    parallel_chunks(T, nt, [&](int64_t ta, int64_t tb, int) {
        for (int64_t t = ta; t < tb; t++) {
            const std::vector<Piece>& tp = tpieces[t];
            const int64_t base = tpiece0[t];
            for (size_t ql = 0; ql < tp.size(); ql++) {
                if (ql + 16 < tp.size()) {   // rows of a group are far apart: fetch ahead
                    const Piece& f = tp[ql + 16];
                    __builtin_prefetch(&s.ptr[f.row]);
                    __builtin_prefetch(&s.col[s.ptr[f.row] + f.k0]);
                    __builtin_prefetch(&s.val_pos[s.ptr[f.row] + f.k0]);
                }
                const int64_t q = base + int64_t(ql);
                const Piece& pc = tp[ql];
                const int32_t r = pc.row;
                const bool last = pc.k0 + pc.len == s.ptr[r + 1] - s.ptr[r];
                p.pc_slot[q] = slot[r];
                p.pc_len[q] = pc.len;
                p.pc_export[q] = last ? xslot[r] : -1;
                p.pc_diag_pos[q] = last ? s.diag_pos[r] : -1;
                for (int32_t k = 0; k < W; k++) {
                    if (k >= pc.len) {
                        p.pc_src[q * W + k] = -1;   // padding: the zero slot, set below
                        p.pc_val_pos[q * W + k] = -1;
                        continue;
                    }
                    const int64_t e = s.ptr[r] + pc.k0 + k, c = s.col[e];
                    p.pc_src[q * W + k] = tile[c] == t ? slot[c] : -(xslot[c] + 1);
                    p.pc_val_pos[q * W + k] = s.val_pos[e];
                }
            }
        }
    }, 1);
    std::vector<std::vector<Piece>>().swap(tpieces);

    // 6. imports: the values each tile reads from other tiles, numbered per
    //    tile in order of first use; they follow the rows in LDS
    std::vector<std::vector<int32_t>> timports{size_t(T)};
    p.group_need.assign(size_t(p.ngroups), 0);
    {
        const int32_t mtr = int32_t(p.max_tile_rows);
        std::vector<std::vector<int32_t>> local{size_t(nt)};
        parallel_chunks(T, nt, [&](int64_t ta, int64_t tb, int th) {
            std::vector<int32_t>& loc = local[size_t(th)];   // exchange slot -> import of the current tile
            loc.assign(size_t(p.nexport), -1);
            for (int64_t t = ta; t < tb; t++) {
                std::vector<int32_t>& imp = timports[t];
                for (int32_t gi = p.tile_group_begin[t]; gi < p.tile_group_begin[t + 1]; gi++) {
                    for (int64_t q = p.group_piece_begin[gi]; q < p.group_piece_begin[gi + 1]; q++)
                        for (int32_t k = 0; k < p.pc_len[q]; k++) {
                            int32_t& src = p.pc_src[q * W + k];
                            if (src >= 0) continue;
                            const int32_t xs = -(src + 1);
                            if (loc[xs] < 0) {
                                loc[xs] = int32_t(imp.size());
                                imp.push_back(xs);
                            }
                            src = mtr + loc[xs];
                        }
                    p.group_need[gi] = int32_t(imp.size());
                }
                for (int32_t xs : imp) loc[xs] = -1;
            }
        }, 1);
    }
    p.tile_import_begin.assign(T + 1, 0);
    for (int64_t t = 0; t < T; t++) {
        p.tile_import_begin[t + 1] = p.tile_import_begin[t] + int32_t(timports[t].size());
        p.max_tile_imports = std::max<int64_t>(p.max_tile_imports, int64_t(timports[t].size()));
    }
    p.nimports = p.tile_import_begin[T];
    p.imports.resize(size_t(p.nimports));
    parallel_chunks(T, nt, [&](int64_t ta, int64_t tb, int) {
        for (int64_t t = ta; t < tb; t++)
            std::copy(timports[t].begin(), timports[t].end(), p.imports.begin() + p.tile_import_begin[t]);
    }, 1);
    // padding entries read the zero slot after the imports with value 0:
    // rr - 0 * (+0) == rr exactly for every rr, so padding changes nothing
    if (p.max_tile_rows + p.max_tile_imports + 1 > 65535) throw std::runtime_error("sweep tile exceeds LDS index range");
    const int32_t zero_slot = int32_t(p.max_tile_rows + p.max_tile_imports);
    parallel_chunks(p.npieces, nt, [&](int64_t q0, int64_t q1, int) {
        for (int64_t q = q0; q < q1; q++)
            for (int32_t k = p.pc_len[q]; k < W; k++) p.pc_src[q * W + k] = zero_slot;
    });
    return p;
}

// ---------------------------------------------------------------------------
// validation and CPU emulation
// ---------------------------------------------------------------------------
std::string validate_sweep_plan(const SweepPlan& p, const SweepRows& s)
{
    const int64_t n = p.n;
    const int W = p.width;
    if (int64_t(p.rows.size()) != n) return "rows size";
    if (int64_t(p.row_slot.size()) != n) return "row_slot size";
    std::vector<int32_t> tile_of(n, -1), slot_of(n, -1), x_of(n, -1), row_at(n, -1);
    for (int64_t t = 0; t < p.ntiles; t++) {
        const int32_t rb = p.tile_row_begin[t], nrows = p.tile_row_begin[t + 1] - rb;
        if (nrows > p.max_tile_rows) return "tile rows";
        for (int32_t q = rb; q < rb + nrows; q++) {
            int32_t r = p.rows[q];
            if (r < 0 || r >= n || tile_of[r] != -1) return "row listed twice or out of range";
            const int32_t sl = p.row_slot[q];
            if (sl < 0 || sl >= nrows || row_at[rb + sl] != -1) return "slots are not a permutation";
            tile_of[r] = int32_t(t);
            slot_of[r] = sl;
            row_at[rb + sl] = r;
        }
    }
    for (int64_t r = 0; r < n; r++)
        if (tile_of[r] < 0) return "row missing";
    if (p.group_piece_begin.front() != 0 || p.group_piece_begin.back() != p.npieces) return "piece range";
    // walk each tile's groups: pieces of a row in entry order, dependencies final
    std::vector<int64_t> done_k(n, 0);        // entries of the row consumed so far
    std::vector<int64_t> final_group(n, -1);  // group of the row's last piece
    std::vector<int64_t> piece_group(n, -1);  // group of the row's latest piece
    for (int64_t t = 0; t < p.ntiles; t++) {
        const int32_t rb = p.tile_row_begin[t], nrows = p.tile_row_begin[t + 1] - rb;
        int32_t xcount = 0;   // exports of the tile so far
        for (int32_t gi = p.tile_group_begin[t]; gi < p.tile_group_begin[t + 1]; gi++) {
            const int64_t m = p.group_piece_begin[gi + 1] - p.group_piece_begin[gi];
            if (m < 1 || m > 64) return "group size";
            if (p.group_xbefore[gi] != xcount) return "exports before group";
            for (int64_t q = p.group_piece_begin[gi]; q < p.group_piece_begin[gi + 1]; q++) {
                const int32_t sl = p.pc_slot[q];
                if (sl < 0 || sl >= nrows) return "piece slot range";
                const int32_t r = row_at[rb + sl];
                if (piece_group[r] == gi) return "two pieces of a row in one group";
                if (final_group[r] >= 0) return "piece after the row's last piece";
                piece_group[r] = gi;
                const int64_t len = s.ptr[r + 1] - s.ptr[r];
                if (p.pc_len[q] < 0 || p.pc_len[q] > W || (p.pc_len[q] == 0 && len != 0)) return "piece length";
                for (int32_t k = 0; k < W; k++) {
                    const int32_t src = p.pc_src[q * W + k];
                    if (k >= p.pc_len[q]) {
                        if (p.pc_val_pos[q * W + k] != -1) return "padding value";
                        if (src != p.max_tile_rows + p.max_tile_imports) return "padding source";
                        continue;
                    }
                    if (src < 0 || src >= p.max_tile_rows + p.max_tile_imports) return "LDS index range";
                    const int64_t e = s.ptr[r] + done_k[r] + k;
                    if (e >= s.ptr[r + 1]) return "piece past the row's end";
                    const int64_t c = s.col[e];
                    if (p.pc_val_pos[q * W + k] != s.val_pos[e]) return "entry value position";
                    if (tile_of[c] == t) {
                        if (src != slot_of[c]) return "intra-tile source";
                        if (final_group[c] < 0 || final_group[c] >= gi) return "intra-tile dependency not final";
                    } else {
                        if (tile_of[c] >= t) return "dependency on a later tile";
                        const int32_t imp = src - int32_t(p.max_tile_rows);
                        if (imp < 0 || p.tile_import_begin[t] + imp >= p.tile_import_begin[t + 1]) return "import range";
                        if (x_of[c] < 0 || p.imports[p.tile_import_begin[t] + imp] != x_of[c]) return "missing export";
                        if (imp >= p.group_need[gi]) return "import not counted before its first use";
                    }
                }
                done_k[r] += p.pc_len[q];
                const bool last = done_k[r] == len;
                if (p.pc_diag_pos[q] != (last ? s.diag_pos[r] : -1)) return "diag pos";
                if (last) {
                    final_group[r] = gi;
                    x_of[r] = p.pc_export[q];
                    if (x_of[r] >= 0) {
                        if (x_of[r] != p.tile_export_begin[t] + xcount) return "export slots out of finish order";
                        if (p.xlist[x_of[r]] != slot_of[r]) return "export list slot";
                        xcount++;
                    }
                } else if (p.pc_export[q] != -1) {
                    return "export before the last piece";
                }
            }
        }
    }
    for (int64_t r = 0; r < n; r++)
        if (final_group[r] < 0) return "row not finished";
    for (int64_t t = 0; t < p.ntiles; t++)
        if (p.tile_export_begin[t + 1] - p.tile_export_begin[t] > p.max_tile_exports) return "tile exports";
    return "";
}

template <typename FV, typename VT>
void emulate_sweep(const SweepPlan& p, const FV* f, const VT* v, VT* y)
{
    const int W = p.width;
    std::vector<VT> lds(p.max_tile_rows + p.max_tile_imports + 1, VT(0)), xbuf(p.nexport);
    for (int64_t t = 0; t < p.ntiles; t++) {
        const int32_t rb = p.tile_row_begin[t], re = p.tile_row_begin[t + 1];
        for (int32_t q = rb; q < re; q++) lds[p.row_slot[q]] = v[p.rows[q]];
        for (int32_t i = p.tile_import_begin[t]; i < p.tile_import_begin[t + 1]; i++)
            lds[p.max_tile_rows + i - p.tile_import_begin[t]] = xbuf[p.imports[i]];
        for (int32_t gi = p.tile_group_begin[t]; gi < p.tile_group_begin[t + 1]; gi++)
            for (int64_t q = p.group_piece_begin[gi]; q < p.group_piece_begin[gi + 1]; q++) {
                VT rr = lds[p.pc_slot[q]];
                for (int32_t k = 0; k < W; k++) {   // padding: 0 * (zero slot)
                    const int64_t vp = p.pc_val_pos[q * W + k];
                    rr -= VT(vp >= 0 ? f[vp] : FV(0)) * lds[p.pc_src[q * W + k]];
                }
                if (p.pc_diag_pos[q] >= 0) rr = VT(rr * VT(f[p.pc_diag_pos[q]]));
                lds[p.pc_slot[q]] = rr;
                if (p.pc_export[q] >= 0) xbuf[p.pc_export[q]] = rr;
            }
        for (int32_t q = rb; q < re; q++) y[p.rows[q]] = lds[p.row_slot[q]];
    }
}

template void emulate_sweep<double, double>(const SweepPlan&, const double*, const double*, double*);
template void emulate_sweep<float, double>(const SweepPlan&, const float*, const double*, double*);
template void emulate_sweep<float, float>(const SweepPlan&, const float*, const float*, float*);

}  // namespace m4g

namespace m4g {

// Rows and factor-storage indices of the sweeps are 32-bit.
static void check_32bit(int64_t n, int64_t entries)
{
    if (n >= (int64_t(1) << 31) || entries >= (int64_t(1) << 31))
        throw std::length_error("matrix too large for the sweeps: 2^31 or more rows or stored entries");
}

bool upper_pattern(int64_t n, const int64_t* S, const int64_t* I, PlanVec64& US, PlanVec64& UI, PlanVec32& a_pos)
{
    // two passes over the rows (count, then fill), each split over the
    // planning threads; a U row must start with its diagonal
    check_32bit(n, n ? S[n] : 0);
    const int nt = n >= 65536 ? plan_threads() : 1;
    US.resize(n + 1);
    US[0] = 0;
    int bad = 0;
    parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int) {
        for (int64_t r = r0; r < r1; r++) {
            const int64_t* b = I + S[r];
            const int64_t* e = I + S[r + 1];
            const int64_t* f = std::lower_bound(b, e, r);   // columns ascend
            if (f == e || *f != r) {
                __atomic_store_n(&bad, 1, __ATOMIC_RELAXED);
                return;
            }
            US[r + 1] = e - f;
        }
    });
    if (bad) return false;
    for (int64_t r = 0; r < n; r++) US[r + 1] += US[r];
    UI.resize(size_t(US[n]));
    a_pos.resize(size_t(US[n]));
    parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int) {
        for (int64_t r = r0; r < r1; r++) {
            const int64_t first = S[r + 1] - (US[r + 1] - US[r]);
            for (int64_t k = 0; k < US[r + 1] - US[r]; k++) {
                UI[US[r] + k] = I[first + k];
                a_pos[US[r] + k] = int32_t(first + k);
            }
        }
    });
    return true;
}

bool ic0_rows(int64_t n, const PlanVec64& US, const PlanVec64& UI, SweepRows& lower, SweepRows& upper)
{
    check_32bit(n, n ? US[n] : 0);
    // upper sweep on U: diagonal first, then entries in column order (on a
    // second host thread; the lower rows below do not depend on it)
    bool upper_ok = true;
    std::thread upper_thread([&] {
        for (int64_t r = 0; r < n; r++)
            if (US[r] == US[r + 1] || UI[US[r]] != r) {
                upper_ok = false;
                return;
            }
        // row r keeps U's row without its diagonal: entries US[r] + 1 ..
        upper.n = n;
        upper.lower = false;
        upper.ptr.resize(n + 1);
        upper.diag_pos.resize(n);
        upper.col.resize(UI.size() - size_t(n));
        upper.val_pos.resize(UI.size() - size_t(n));
        const int nt = n >= 65536 ? std::max(1, plan_threads() - 1) : 1;
        parallel_chunks(n + 1, nt, [&](int64_t r0, int64_t r1, int) {
            for (int64_t r = r0; r < r1; r++) {
                upper.ptr[r] = US[r] - r;
                if (r == n) continue;
                upper.diag_pos[r] = int32_t(US[r]);
                for (int64_t p = US[r] + 1; p < US[r + 1]; p++) {
                    upper.col[p - r - 1] = int32_t(UI[p]);
                    upper.val_pos[p - r - 1] = int32_t(p);
                }
            }
        });
    });
    struct Join {
        std::thread& t;
        ~Join() { if (t.joinable()) t.join(); }
    } join{upper_thread};
    // lower sweep on L = trans(U): row i holds U's column i (rows k < i
    // ascending), inverted diagonal last.  Split over threads by ranges of
    // L rows when U is banded enough: the thread of L rows [c0, c1) scans the
    // U rows [c0 - band, c1), in ascending order, so each L row is filled in
    // the same order as by one pass over all U rows.
    lower.n = n;
    lower.lower = true;
    lower.ptr.resize(n + 1);
    lower.diag_pos.resize(n);
    int64_t band = 0;   // largest column - row of U
    for (int64_t r = 0; r < n; r++)
        if (US[r + 1] - US[r] > 1) band = std::max(band, UI[US[r + 1] - 1] - r);
    const int nt = n >= 65536 && band * plan_threads() <= n ? std::max(1, plan_threads() - 1) : 1;
    // entries per L row (U's column counts), over the same row ranges
    lower.ptr[0] = 0;
    parallel_chunks(n, nt, [&](int64_t c0, int64_t c1, int) {
        std::fill(lower.ptr.begin() + c0 + 1, lower.ptr.begin() + c1 + 1, 0);
        for (int64_t r = std::max<int64_t>(0, c0 - band); r < c1; r++) {
            const int64_t* b = UI.data() + US[r] + 1;
            const int64_t* e = UI.data() + US[r + 1];
            for (const int64_t* q = std::lower_bound(b, e, c0); q < e && *q < c1; q++) lower.ptr[*q + 1]++;
        }
    });
    for (int64_t r = 0; r < n; r++) lower.ptr[r + 1] += lower.ptr[r];
    lower.col.resize(size_t(lower.ptr[n]));
    lower.val_pos.resize(size_t(lower.ptr[n]));
    parallel_chunks(n, nt, [&](int64_t c0, int64_t c1, int) {
        std::vector<int64_t> fill(lower.ptr.begin() + c0, lower.ptr.begin() + c1);
        for (int64_t r = std::max<int64_t>(0, c0 - band); r < c1; r++) {
            if (r >= c0) lower.diag_pos[r] = int32_t(US[r]);
            // entries of U row r with a column in [c0, c1) (columns ascend)
            const int64_t* b = UI.data() + US[r] + 1;
            const int64_t* e = UI.data() + US[r + 1];
            for (const int64_t* q = std::lower_bound(b, e, c0); q < e && *q < c1; q++) {
                const int64_t d = fill[*q - c0]++;
                lower.col[d] = int32_t(r);
                lower.val_pos[d] = int32_t(q - UI.data());
            }
        }
    });
    upper_thread.join();
    return upper_ok;
}

bool ilu0_rows(int64_t n, const int64_t* S, const int64_t* I, SweepRows& lower, SweepRows& upper,
               PlanVec64& diag_pos)
{
    // L: the strictly lower entries of each row, U: the strictly upper ones
    // (diagonal apart); counted, then filled, over the planning threads
    check_32bit(n, n ? S[n] : 0);
    lower.n = upper.n = n;
    lower.lower = true;
    upper.lower = false;
    lower.ptr.resize(n + 1);
    upper.ptr.resize(n + 1);
    lower.diag_pos.assign(n, -1);
    upper.diag_pos.resize(n);
    diag_pos.resize(n);
    lower.ptr[0] = upper.ptr[0] = 0;
    const int nt = n >= 65536 ? plan_threads() : 1;
    int bad = 0;
    parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int) {
        for (int64_t r = r0; r < r1; r++) {
            const int64_t* b = I + S[r];
            const int64_t* e = I + S[r + 1];
            const int64_t* f = std::lower_bound(b, e, r);   // columns ascend
            if (f == e || *f != r) {
                __atomic_store_n(&bad, 1, __ATOMIC_RELAXED);
                return;
            }
            diag_pos[r] = f - I;
            upper.diag_pos[r] = int32_t(f - I);
            lower.ptr[r + 1] = f - b;
            upper.ptr[r + 1] = e - f - 1;
        }
    });
    if (bad) return false;
    for (int64_t r = 0; r < n; r++) {
        lower.ptr[r + 1] += lower.ptr[r];
        upper.ptr[r + 1] += upper.ptr[r];
    }
    lower.col.resize(size_t(lower.ptr[n]));
    lower.val_pos.resize(size_t(lower.ptr[n]));
    upper.col.resize(size_t(upper.ptr[n]));
    upper.val_pos.resize(size_t(upper.ptr[n]));
    parallel_chunks(n, nt, [&](int64_t r0, int64_t r1, int) {
        for (int64_t r = r0; r < r1; r++) {
            int64_t lq = lower.ptr[r], uq = upper.ptr[r];
            for (int64_t p = S[r]; p < diag_pos[r]; p++, lq++) {
                lower.col[lq] = int32_t(I[p]);
                lower.val_pos[lq] = int32_t(p);
            }
            for (int64_t p = diag_pos[r] + 1; p < S[r + 1]; p++, uq++) {
                upper.col[uq] = int32_t(I[p]);
                upper.val_pos[uq] = int32_t(p);
            }
        }
    });
    return true;
}

bool ilu0_adjoint_rows(int64_t n, const int64_t* S, const int64_t* I, SweepRows& lower, SweepRows& upper)
{
    check_32bit(n, n ? S[n] : 0);
    std::vector<int64_t> dpos(n, -1);
    lower.n = upper.n = n;
    lower.lower = true;
    upper.lower = false;
    lower.ptr.assign(n + 1, 0);
    upper.ptr.assign(n + 1, 0);
    lower.diag_pos.assign(n, -1);
    upper.diag_pos.assign(n, -1);
    for (int64_t r = 0; r < n; r++)
        for (int64_t p = S[r]; p < S[r + 1]; p++) {
            if (I[p] == r) dpos[r] = p;
            else if (I[p] > r) lower.ptr[I[p] + 1]++;   // U^T row I[p] gets (I[p], r)
            else upper.ptr[I[p] + 1]++;                 // L^T row I[p] gets (I[p], r)
        }
    for (int64_t r = 0; r < n; r++) {
        if (dpos[r] < 0) return false;
        lower.ptr[r + 1] += lower.ptr[r];
        upper.ptr[r + 1] += upper.ptr[r];
        lower.diag_pos[r] = dpos[r];
    }
    lower.col.assign(lower.ptr[n], 0);
    lower.val_pos.assign(lower.ptr[n], 0);
    upper.col.assign(upper.ptr[n], 0);
    upper.val_pos.assign(upper.ptr[n], 0);
    std::vector<int64_t> lf(lower.ptr.begin(), lower.ptr.end() - 1);
    // ascending r for U^T rows (scatter order of lower_trisolve.hpp:155-170)
    for (int64_t r = 0; r < n; r++)
        for (int64_t p = S[r]; p < S[r + 1]; p++)
            if (I[p] > r) {
                int64_t d = lf[I[p]]++;
                lower.col[d] = r;
                lower.val_pos[d] = p;
            }
    // descending r for L^T rows (scatter order of upper_trisolve.hpp:131-149)
    std::vector<int64_t> uf(upper.ptr.begin(), upper.ptr.end() - 1);
    for (int64_t r = n - 1; r >= 0; r--)
        for (int64_t p = S[r]; p < S[r + 1]; p++)
            if (I[p] < r) {
                int64_t d = uf[I[p]]++;
                upper.col[d] = r;
                upper.val_pos[d] = p;
            }
    return true;
}

}  // namespace m4g
