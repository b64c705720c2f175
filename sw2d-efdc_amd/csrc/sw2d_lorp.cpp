// sw2d_lorp.cpp -- the domain decomposition file of the coupled EFDC run
// (EFDC/input/LORP.INP).  The reference runs its partitioner on CELL.INP
// (SW2D-EFDC.cu:4480-4487: "Gorp CELL.INP sc <N>", N = setup.dat's number
// of processors) and renames LORP_sc_<N>_{h,v,r}.INP by setup.dat's LORP
// type to EFDC/input/LORP.INP (4491-4507).  The partitioner is Gorp
// (Gorp.h, Gorp.cpp, GorpMain.cpp); this file computes the same partition
// and writes the same file, from the cell types directly.
//
// The problem.  Cut the grid into nh columns of tiles (widths wx) and nv rows
// of tiles (heights wy, from the south).  A tile's cost models one EFDC
// rank's work (Gorp.h:219-236): zero without lake (5) and boundary (9)
// cells, else a ghost-perimeter term, the lake cells scaled to the tile area
// with a 2-cell ghost ring, and the boundary cells, with fixed weights.  A
// layout's score is its largest tile cost plus 1e10 x |1 + active - N| when
// the number of tiles with a positive cost differs from N.
//
// The search (Gorp.h:461-646).  Start from cuts that give every band about
// the same number of lake cells (rows from the south, then columns, targets
// computed in single precision); then, for steps d = 32, 16, ..., 1, move
// the single cut by +d or -d that lowers the score most (columns' cuts
// first, + before -, the first of equal scores kept), repeat while one
// helps, and restart from d = 32 while any step moved.  Type h searches N
// columns of tiles, type v N rows, type r every nv = 1..N+2 with
// nh = N/nv..N+2 and nh nv >= N, keeping the first lowest score.
//
// Where the reference rescans a tile's cells for every candidate move, the
// counts here come from summed-area tables and a candidate is scored from the
// tiles it changes (O(nh + nv) per move).  The scores are the same doubles
// (integer counts, the same operations in the same order), so every
// comparison, and the file, comes out the same.
//
// File (Gorp.h:674-739): its original name, a description line, the header
// of the summary values, NH NV NActive Cost and the five weights (C++ stream
// default formatting, i.e. %g with 6 digits), the tile widths and heights
// each plus the 4-cell ghost zone, the ids of the tiles with a positive cost
// (rows of tiles from the south, west to east), and the id -> rank table.
// The texts are the file's format, kept as the reference writes them.
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "sw2d_io.hpp"

namespace sw2dio {

namespace {

// cost weights of the reference's model; the area terms are zero there
constexpr double kWeightTotalArea = 0.00 / 203 / 131;
constexpr double kWeightTileArea = 0.00 / 44 / 25;
constexpr double kWeightGhost = 0.001 / (4 * 44 + 4 * 25);
constexpr double kWeightLake = 0.007 / 44 / 25;
constexpr double kWeightLand = 0.00 / 44 / 25;
constexpr double kWeightBoundary = 0.003 / 44 / 25;
constexpr double kInvalid = 1e50;   // cost of a tile outside the grid or empty
constexpr double kPenalty = 1e10;   // per rank of difference in the active count
const int kGhostCells = 4;      // added to every width in the file
const int kFirstStep = 32;

// Summed-area tables of the lake, boundary and land cells; y counts rows
// from the south, as the tile heights do.
class Counts {
public:
    Counts(const std::vector<unsigned char> &types, int cols, int rows) : W(cols), H(rows)
    {
        const size_t n = (size_t)(W + 1) * (H + 1);
        lake.assign(n, 0); bnd.assign(n, 0); land.assign(n, 0);
        for (int y = 0; y < H; y++) {
            const unsigned char *row = &types[(size_t)(H - 1 - y) * W];  // types: north row first
            for (int x = 0; x < W; x++) {
                const size_t o = at(x + 1, y + 1);
                const int t = row[x];
                lake[o] = (t == 5) + lake[at(x, y + 1)] + lake[at(x + 1, y)] - lake[at(x, y)];
                bnd[o] = (t == 9) + bnd[at(x, y + 1)] + bnd[at(x + 1, y)] - bnd[at(x, y)];
                land[o] = (t == 0) + land[at(x, y + 1)] + land[at(x + 1, y)] - land[at(x, y)];
            }
        }
    }
    int W, H;
    long long lake_in(int x, int y, int l, int w) const { return box(lake, x, y, l, w); }
    long long bnd_in(int x, int y, int l, int w) const { return box(bnd, x, y, l, w); }
    long long land_in(int x, int y, int l, int w) const { return box(land, x, y, l, w); }

private:
    std::vector<long long> lake, bnd, land;
    size_t at(int x, int y) const { return (size_t)y * (W + 1) + x; }
    long long box(const std::vector<long long> &t, int x, int y, int l, int w) const
    {
        return t[at(x + l, y + w)] - t[at(x, y + w)] - t[at(x + l, y)] + t[at(x, y)];
    }
};

// cost of the tile of l columns and w rows at (x, y) (Gorp.h:219-255)
double tile_cost(const Counts &m, int x, int y, int l, int w)
{
    if (y < 0 || y + w > m.H || x < 0 || x + l > m.W || l <= 0 || w <= 0) return kInvalid;
    const double lake = (double)m.lake_in(x, y, l, w);
    const double bnd = (double)m.bnd_in(x, y, l, w);
    const double land = (double)m.land_in(x, y, l, w);
    if (lake == 0 && bnd == 0) return 0.0;
    const double ring_area = (double)((w + 4) * (l + 4));
    const double lake_est = lake / (double)(w * l) * ring_area;
    return kWeightGhost * (double)(2 * w + 2 * l) + kWeightLake * lake_est + kWeightBoundary * bnd +
           kWeightLand * land;
}

struct Layout {
    std::vector<int> wx, wy;  // tile widths west to east, heights south to north
    double score = 0;         // largest cost (+ the active-count penalty)
};

class Search {
public:
    Search(const Counts &counts, int n) : m(counts), target(n) {}

    // largest tile cost and the number of tiles with a positive cost
    // (Gorp.h:296-312)
    void stats(const std::vector<int> &wx, const std::vector<int> &wy, double &maxc, int &active) const
    {
        maxc = 0;
        active = 0;
        double area = 0;
        int y = 0;
        for (int w : wy) {
            int x = 0;
            for (int l : wx) {
                const double c = tile_cost(m, x, y, l, w);
                if (c > 0) {
                    active++;
                    if (maxc < c) maxc = c;
                    area += l * w;
                }
                x += l;
            }
            y += w;
        }
        maxc += area * kWeightTotalArea;
    }

    double score(const std::vector<int> &wx, const std::vector<int> &wy) const
    {
        double c;
        int active;
        stats(wx, wy, c, active);
        if (active != target) c += kPenalty * abs(1 + active - target);
        return c;
    }

    // bands of about equal lake-cell counts (Gorp.h:493-539)
    Layout start(int nh, int nv) const
    {
        long long lake = m.lake_in(0, 0, m.W, m.H);
        Layout L;
        L.wx = bands(nh, (int)lake, m.W, [&](int i) { return m.lake_in(i, 0, 1, m.H); });
        L.wy = bands(nv, (int)lake, m.H, [&](int i) { return m.lake_in(0, i, m.W, 1); });
        return L;
    }

    // the best single cut move by +-d (Gorp.h:579-635); true when one helped.
    // A move of cut i changes only the tiles on either side of it, so a
    // candidate's score takes the other tiles' largest cost and active count
    // from per-band maxima and counts of the current layout (the same doubles
    // as a full evaluation: the total-area weight is zero, so that term adds
    // +0.0 to a score whichever tiles are active).
    bool improve(Layout &L, int d) const
    {
        static_assert(kWeightTotalArea == 0.0, "incremental scores leave out the area term");
        const int nh = (int)L.wx.size(), nv = (int)L.wy.size();
        std::vector<int> X(nh + 1, 0), Y(nv + 1, 0);
        for (int i = 0; i < nh; i++) X[i + 1] = X[i] + L.wx[i];
        for (int r = 0; r < nv; r++) Y[r + 1] = Y[r] + L.wy[r];
        // current tile costs, their maxima and active counts per column and row
        std::vector<double> cmax(nh, 0.0), rmax(nv, 0.0);
        std::vector<int> cact(nh, 0), ract(nv, 0);
        int act = 0;
        double top = 0;
        for (int r = 0; r < nv; r++)
            for (int i = 0; i < nh; i++) {
                const double c = tile_cost(m, X[i], Y[r], L.wx[i], L.wy[r]);
                if (c > 0) {
                    act++; cact[i]++; ract[r]++;
                    if (cmax[i] < c) cmax[i] = c;
                    if (rmax[r] < c) rmax[r] = c;
                    if (top < c) top = c;
                }
            }
        auto penalised = [&](double c, int a) { return a != target ? c + kPenalty * abs(1 + a - target) : c; };
        double best = penalised(top, act);
        // the largest band maximum outside bands i and i+1
        auto others = [](const std::vector<double> &mx, int i) {
            double o = 0;
            for (int k = 0; k < (int)mx.size(); k++)
                if (k != i && k != i + 1 && o < mx[k]) o = mx[k];
            return o;
        };
        int axis = -1, cut = 0, step = 0;
        for (int ax = 0; ax < 2; ax++) {
            const std::vector<int> &w = ax == 0 ? L.wx : L.wy;
            const std::vector<double> &mx = ax == 0 ? cmax : rmax;
            const std::vector<int> &na = ax == 0 ? cact : ract;
            const std::vector<int> &P = ax == 0 ? X : Y;
            const int nb = (int)w.size(), nc = ax == 0 ? nv : nh;
            for (int i = 0; i + 1 < nb; i++) {
                const double rest = others(mx, i);
                for (int s : {d, -d}) {
                    const int a0 = w[i] + s, a1 = w[i + 1] - s, p1 = P[i] + a0;
                    double c = rest;
                    int a = act - na[i] - na[i + 1];
                    for (int k = 0; k < nc; k++) {
                        const double t0 = ax == 0 ? tile_cost(m, P[i], Y[k], a0, L.wy[k])
                                                  : tile_cost(m, X[k], P[i], L.wx[k], a0);
                        const double t1 = ax == 0 ? tile_cost(m, p1, Y[k], a1, L.wy[k])
                                                  : tile_cost(m, X[k], p1, L.wx[k], a1);
                        if (t0 > 0) { a++; if (c < t0) c = t0; }
                        if (t1 > 0) { a++; if (c < t1) c = t1; }
                    }
                    const double sc = penalised(c, a);
                    if (sc < best) { best = sc; axis = ax; cut = i; step = s; }
                }
            }
        }
        if (axis < 0) return false;
        std::vector<int> &w = axis == 0 ? L.wx : L.wy;
        w[cut] += step;
        w[cut + 1] -= step;
        return true;
    }

    // Gorp.h:638-651 (findBestLayoutNear)
    Layout refine(int nh, int nv) const
    {
        Layout L = start(nh, nv);
        bool moved = true;
        while (moved) {
            moved = false;
            for (int d = kFirstStep; d > 0; d /= 2)
                while (improve(L, d)) moved = true;
        }
        L.score = score(L.wx, L.wy);
        return L;
    }

private:
    const Counts &m;
    int target;

    template <typename F> static std::vector<int> bands(int n, int lake, int len, F count_at)
    {
        std::vector<int> w(n);
        int pos = 0, prev = 0, got = 0, sum = 0;
        for (int i = 0; i + 1 < n; i++) {
            const float f = (float)(i + 1) * (float)lake / (float)n;
            const int goal = (int)f;
            while (got < goal && pos < len) got += (int)count_at(pos++);
            w[i] = pos - prev;
            sum += w[i];
            prev = pos;
        }
        w[n - 1] = len - sum;
        return w;
    }
};

}  // namespace

bool write_lorp(const std::string &path, const std::vector<unsigned char> &types, int cols, int rows,
                int nproc, int kind, std::string *err)
{
    auto fail = [&](const std::string &m) {
        if (err) *err = m;
        return false;
    };
    if (cols < 1 || rows < 1 || types.size() != (size_t)cols * rows) return fail("LORP: bad grid");
    if (nproc < 1) return fail("LORP: the number of processors must be at least 1");
    const Counts m(types, cols, rows);
    const Search s(m, nproc);
    Layout best;
    char tag = 'r';
    if (kind == 1 || kind == 2) {
        tag = kind == 1 ? 'h' : 'v';
        best = kind == 1 ? s.refine(nproc, 1) : s.refine(1, nproc);
    } else {
        bool any = false;
        for (int nv = 1; nv <= nproc + 2; nv++)
            for (int nh = nproc / nv; nh <= nproc + 2; nh++) {
                if (nv * nh < nproc) continue;
                Layout L = s.refine(nh, nv);
                if (!any || L.score < best.score) best = L;
                any = true;
            }
    }
    // the file states the chosen layout's cost without the penalty
    double cost;
    int active;
    s.stats(best.wx, best.wy, cost, active);
    std::vector<int> ids;
    {
        int y = 0, id = 0;
        for (int w : best.wy) {
            int x = 0;
            for (int l : best.wx) {
                if (tile_cost(m, x, y, l, w) > 0) ids.push_back(id);
                x += l;
                id++;
            }
            y += w;
        }
    }
    FILE *fp = fopen(path.c_str(), "w");
    if (!fp) return fail("cannot write " + path);
    const int nh = (int)best.wx.size(), nv = (int)best.wy.size();
    fprintf(fp, "LORP_sc_%d_%c.INP\n", nproc, tag);
    fprintf(fp, " LORP input data used for decomposing domain into optimal configuration with %d processors\n",
            nproc);
    fprintf(fp, " NH NV NActive Cost TotalAreaWeight TileAreaWeight TileGhostWeight TileWaterWeight "
                "TileNearLandWeight\n");
    fprintf(fp, " %d %d %d %g %g %g %g %g %g\n", nh, nv, active, cost, kWeightTotalArea, kWeightTileArea,
            kWeightGhost, kWeightLake, kWeightBoundary);
    fprintf(fp, "*   IC:    an array of NPARTX size setting x-width for each partition\n");
    for (int l : best.wx) fprintf(fp, "%d\n", l + kGhostCells);
    fprintf(fp, "*   JC:    an array of NPARTY size setting y-height for each partition\n");
    for (int w : best.wy) fprintf(fp, "%d\n", w + kGhostCells);
    fprintf(fp, "*   List of active partition IDs, 0-based\n");
    for (int id : ids) fprintf(fp, "%d\n", id);
    fprintf(fp, "*   Lookup table to go from your partition id to node id (so has full N entries and "
                "some are -1, 0-based)\n");
    for (int id = 0, k = 0; id < nh * nv; id++) {
        const bool on = k < (int)ids.size() && ids[k] == id;
        fprintf(fp, "%d\n", on ? k : -1);
        k += on;
    }
    const bool ok = !ferror(fp);
    if (fclose(fp) != 0 || !ok) return fail("cannot write " + path);
    return true;
}

bool read_cell_types(const std::string &path, std::vector<unsigned char> &types, int &cols, int &rows,
                     std::string *err)
{
    auto fail = [&](const std::string &m) {
        if (err) *err = m;
        return false;
    };
    FILE *fp = fopen(path.c_str(), "r");
    if (!fp) return fail("cannot read " + path);
    std::vector<std::string> lines;
    {
        std::string cur;
        int ch;
        while ((ch = fgetc(fp)) != EOF) {
            if (ch == '\n') { lines.push_back(cur); cur.clear(); }
            else cur += (char)ch;
        }
        if (!cur.empty()) lines.push_back(cur);
        fclose(fp);
    }
    char name[256];
    if (lines.empty() || sscanf(lines[0].c_str(), "C %255s file, %d columns and %d rows", name, &cols, &rows) != 3 ||
        cols < 1 || rows < 1)
        return fail(path + ": no \"C <name> file, <cols> columns and <rows> rows\" title");
    types.assign((size_t)cols * rows, 0);
    // blocks of rows from the north after the four header lines; the type
    // digits start at character 6 of a line
    size_t k = 4;
    int done = 0;
    while (done < cols) {
        int width = -1;
        for (int r = 0; r < rows; r++, k++) {
            if (k >= lines.size()) return fail(path + ": ends inside a block of rows");
            const std::string &l = lines[k];
            int n = 0;
            for (size_t p = 5; p < l.size() && l[p] >= '0' && l[p] <= '9'; p++, n++)
                if (done + n < cols) types[(size_t)r * cols + done + n] = (unsigned char)(l[p] - '0');
            if (width < 0) width = n;
        }
        if (width <= 0) return fail(path + ": a block of rows without cell types");
        done += width;
    }
    return true;
}

}  // namespace sw2dio
