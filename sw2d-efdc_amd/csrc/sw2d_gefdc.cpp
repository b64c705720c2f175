// sw2d_gefdc.cpp -- the EFDC grid-generator inputs the coupled run writes
// before stepping (make_gefdc_inputs, SW2D-EFDC.cu:2831-3680, called at 4472
// with mask.dat and the DEM after lake_catchment_smooth): CELL.INP (cell
// types of the lake grid), gridext.inp (cell corner coordinates),
// depdat.inp (lake-cell centres and bed elevations) and gefdc.inp (the grid
// generator's control cards, from the case's grid_make_input.dat).
//
// Cell types.  A cell's contour flag is set for lake cells (mask 1) and for
// every other cell whose mask value differs from one of its 4-neighbours in
// the grid.  CELL.INP gives lake cells type 5; another cell is 9 when its
// flag is set or differs from a neighbour's flag, else 0.  For the type the
// reference looks at east/west neighbours only inside the block of at most
// 120 columns the cell is printed in (2956-3196: a block's first column has
// no west neighbour, its last no east one; a one-column block looks west),
// and at the north/south neighbours that exist.  The reference leaves the
// contour flag of a lake cell in a grid corner unset (it reads whatever the
// allocation held); here it is set, like every other lake cell's.
//
// Layout.  CELL.INP: a title line, two column-number lines, "C", then the
// blocks of 120 columns, each listing every row from the north with its row
// number (right-aligned, width 2-5 by the row count, then two spaces; rows
// numbered from 100000 on get no number, as in the reference), then "C" and
// the two column-number lines again.  gridext.inp lists every cell of every
// block, block by block, row by row from the north: I, J, x and y of its
// south-west corner.  depdat.inp lists the lake cells in the same order:
// centre x, y and the bed elevation as a float.  The card and line texts are
// the files' format, kept as the reference prints them.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "sw2d_io.hpp"

namespace sw2dio {

namespace {

struct Out {
    FILE *fp = nullptr;
    explicit Out(const std::string &p) { fp = fopen(p.c_str(), "w"); }
    ~Out() { if (fp) fclose(fp); }
    bool close()
    {
        const bool ok = fp && !ferror(fp) && fclose(fp) == 0;
        fp = nullptr;
        return ok;
    }
};

// first column-number line ("C    0        1         2 ..."): the labels
// 1 .. last-1, each right-aligned in ten characters; `nl_after_one` puts the
// line end right after "1"
void tens_line(FILE *fp, int last, bool nl_after_one)
{
    fprintf(fp, "C    0        1%s", nl_after_one ? "\n" : "");
    for (int i = 2; i < last - 1; i++) fprintf(fp, "         %d%s", i, i == last - 2 ? "\n" : "");
}

// second column-number line: "1234567890" once per ten columns
void units_line(FILE *fp, int dec)
{
    fputs("C    ", fp);
    for (int i = 1; i < dec; i++) fprintf(fp, "1234567890%s", i == dec - 1 ? "\n" : "");
}

// the row number column of CELL.INP
void row_label(FILE *fp, int rows, int v)
{
    const int w = rows >= 10000 ? 5 : rows >= 1000 ? 4 : rows >= 10 ? 3 : 2;
    if (rows >= 100000 && v >= 100000) return;
    fprintf(fp, "%*d  ", w, v);
}

}  // namespace

bool read_grid_make_input(const std::string &path, GefdcParams &p, std::string *err)
{
    FILE *f = fopen(path.c_str(), "r");
    if (!f) {
        if (err) *err = "unknown file - grid_make_input.dat";
        return false;
    }
    // the schema of grid_make_input.dat (2944-2991), read as fscanf reads
    // it: section titles, then "NAME: value" lines; a value that does not
    // parse keeps its default (0)
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wformat-security"
    auto lit = [&](const char *s) { if (fscanf(f, s) < 0) return; };
#pragma GCC diagnostic pop
    auto gi = [&](const char *fmt, int &v) { if (fscanf(f, fmt, &v) < 0) return; };
    auto gf = [&](const char *fmt, float &v) { if (fscanf(f, fmt, &v) < 0) return; };
    auto gd = [&](const char *fmt, double &v) { if (fscanf(f, fmt, &v) < 0) return; };
    lit(" SOLUTION ITERATIONS AND SMOOTHING INFORMATION");
    gi(" ITRXM: %d", p.ITRXM); gi(" ITRHM: %d", p.ITRHM); gi(" ITRKM: %d", p.ITRKM);
    gi(" ITRGM: %d", p.ITRGM); gi(" NDEPSM: %d", p.NDEPSM); gi(" NDEPSMF: %d", p.NDEPSMF);
    gf(" DEPMIN: %f", p.DEPMIN); gi(" DDATADJ: %d", p.DDATADJ);
    lit(" PARAMETERS AND CONVERGENCE CRITERIA");
    gf(" RPX: %f", p.RPX); gf(" RPK: %f", p.RPK); gf(" RPH: %f", p.RPH);
    gd(" RSQXM: %le", p.RSQXM); gd(" RSQKM: %le", p.RSQKM); gd(" RSQKIM: %le", p.RSQKIM);
    gd(" RSQHM: %le", p.RSQHM); gd(" RSQHIM: %le", p.RSQHIM); gd(" RSQHJM: %le", p.RSQHJM);
    lit(" COORDINATE SHIFT PARAMETERS");
    gf(" XSHIFT: %f", p.XSHIFT); gf(" YSHIFT: %f", p.YSHIFT); gf(" HSCALE: %f", p.HSCALE);
    gf(" RKJDKI: %f", p.RKJDKI); gf(" ANGORO: %f", p.ANGORO);
    lit(" INTERPOLATION SWITCHES");
    gi(" ISIRKI: %d", p.ISIRKI); gi(" JSIRKI: %d", p.JSIRKI); gi(" ISIHIHJ: %d", p.ISIHIHJ);
    gi(" JSIHIHJ: %d", p.JSIHIHJ);
    lit(" DEPTH INTERPOLATION SWITCHES");
    gi(" ISIDEP: %d", p.ISIDEP); gf(" CDEP: %f", p.CDEP); gf(" RADM: %f", p.RADM);
    gi(" ISIDPTYP: %d", p.ISIDPTYP); gf(" SURFELV: %f", p.SURFELV); gi(" ISVEG: %d", p.ISVEG);
    gi(" NVEGDAT: %d", p.NVEGDAT); gi(" NVEGTYP: %d", p.NVEGTYP); gd(" ZROUGH: %lf", p.ZROUGH);
    gi(" VERBOSE: %d", p.VERBOSE);
    fclose(f);
    return true;
}

std::vector<unsigned char> gefdc_cell_types(int cols, int rows, const std::vector<int> &mask)
{
    const int R = rows, C = cols;
    auto at = [&](int i, int j) { return mask[(size_t)i * C + j]; };
    std::vector<unsigned char> flag((size_t)R * C, 0), type((size_t)R * C, 0);
    for (int i = 0; i < R; i++)
        for (int j = 0; j < C; j++) {
            const int v = at(i, j);
            bool f = v == 1;
            if (!f) {
                f = (i > 0 && at(i - 1, j) != v) || (i < R - 1 && at(i + 1, j) != v) ||
                    (j > 0 && at(i, j - 1) != v) || (j < C - 1 && at(i, j + 1) != v);
            }
            flag[(size_t)i * C + j] = f;
        }
    auto fl = [&](int i, int j) { return flag[(size_t)i * C + j]; };
    const int ncl = 120;
    for (int j_in = 0; j_in < C; j_in += ncl) {
        const int j_end = j_in + ncl < C ? j_in + ncl : C;
        for (int i = 0; i < R; i++)
            for (int j = j_in; j < j_end; j++) {
                const size_t k = (size_t)i * C + j;
                if (at(i, j) == 1) { type[k] = 5; continue; }
                const unsigned char f = fl(i, j);
                bool diff = (i > 0 && fl(i - 1, j) != f) || (i < R - 1 && fl(i + 1, j) != f);
                if (j > j_in && j < j_end - 1)
                    diff = diff || fl(i, j - 1) != f || fl(i, j + 1) != f;
                else if (j >= j_end - 1)
                    diff = diff || (j > 0 && fl(i, j - 1) != f);
                else
                    diff = diff || (j + 1 < C && fl(i, j + 1) != f);
                type[k] = (diff || f) ? 9 : 0;
            }
    }
    return type;
}

bool write_gefdc_inputs(const std::string &dir, int cols, int rows, double res,
                        const std::vector<int> &mask, const std::vector<double> &bed,
                        const GefdcParams &p, int *ndepdat_out, std::string *err)
{
    const std::vector<unsigned char> type = gefdc_cell_types(cols, rows, mask);
    Out cell(dir + "CELL.INP"), grid(dir + "gridext.inp"), dep(dir + "depdat.inp"),
        gef(dir + "gefdc.inp");
    if (!cell.fp || !grid.fp || !dep.fp || !gef.fp) {
        if (err) *err = "cannot write the grid-generator inputs in " + dir;
        return false;
    }
    // blocks of 120 columns; the column-number lines cover one block
    // (2883-2897 of the reference's layout arithmetic)
    const int ncl = 120;
    const int np = cols / ncl + 1;
    int min_dec = 11, width = cols;
    if (np == 1) {
        if (cols / 10 <= 10) min_dec = cols / 10 + 1;
    } else {
        width = ncl;
    }
    const int er = cols - (np - 1) * ncl;
    const int dec = width / 10 + 1;
    fprintf(cell.fp, "C cell.inp file, %d columns and %d rows @\n", cols, rows);
    tens_line(cell.fp, min_dec, min_dec <= 3);
    units_line(cell.fp, dec);
    fputs("C\n", cell.fp);
    int ndepdat = 0;
    for (int sq = 0; sq < np; sq++) {
        const int j_in = sq * ncl;
        const int j_end = (sq == np - 1) ? j_in + er : j_in + ncl;
        for (int i = 0; i < rows; i++) {
            row_label(cell.fp, rows, rows - i);
            for (int j = j_in; j < j_end; j++) {
                fprintf(grid.fp, "%d %d %lf %lf\n", j + 1, rows - i, j * res, (rows - i - 1) * res);
                const size_t k = (size_t)i * cols + j;
                fprintf(cell.fp, "%d%s", type[k], j >= j_end - 1 ? "\n" : "");
                if (type[k] == 5) {
                    ndepdat++;
                    fprintf(dep.fp, "%lf %lf %f\n", (j + (j + 1)) * res / 2,
                            ((rows - i) + ((rows - i) - 1)) * res / 2, (double)(float)bed[k]);
                }
            }
        }
    }
    fputs("C\n", cell.fp);
    units_line(cell.fp, dec);
    tens_line(cell.fp, min_dec, dec <= 2);

    // gefdc.inp (3620-3665)
    FILE *g = gef.fp;
    fputs("C1 TITLE\nC1 (LIMITED TO 80 CHARACTERS)\n'gefdc.inp DHARA-EFDC'\n", g);
    fputs("C2 INTEGER INPUT\nC2 NTYPE  NBPP  IMIN  IMAX  JMIN  JMAX  IC   JC\n", g);
    fprintf(g, "   %d     %d    %d    %d    %d    %d    %d   %d\n", 0, 0, 1, cols, 1, rows, cols, rows);
    fputs("C3 GRAPHICS GRID INFORMATION\nC3 ISGG IGM JGM DXCG DYCG NWTGG\n", g);
    fprintf(g, "   %d   %d  %d  %f   %f   %d\n", 0, 0, 0, res, res, 1);
    fputs("C4 CARTESIAN AND GRAPHICS GRID COORDINATE DATA\nC4 CDLON1 CDLON2 CDLON3 CDLAT1 CDLAT2 CDLAT3\n", g);
    fprintf(g, "   %f     %f     %f     %f     %f     %f\n", 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
    fputs("C5 INTEGER INPUT\nC5 ITRXM ITRHM ITRKM ITRGM NDEPSM NDEPSMF DEPMIN DDATADJ\n", g);
    fprintf(g, "   %d    %d    %d    %d    %d     %d      %f     %d\n", p.ITRXM, p.ITRHM, p.ITRKM,
            p.ITRGM, p.NDEPSM, p.NDEPSMF, (double)p.DEPMIN, p.DDATADJ);
    fputs("C6 REAL INPUT\nC6 RPX RPK RPH RSQXM RSQKM RSQKIM RSQHM RSQHIM RSQHJM\n", g);
    fprintf(g, "   %f  %f  %f  %E    %E    %E     %E    %E     %E\n", (double)p.RPX, (double)p.RPK,
            (double)p.RPH, p.RSQXM, p.RSQKM, p.RSQKIM, p.RSQHM, p.RSQHIM, p.RSQHJM);
    fputs("C7 COORDINATE SHIFT PARAMETERS\nC7 XSHIFT YSHIFT HSCALE RKJDKI ANGORO\n", g);
    fprintf(g, "   %f     %f     %f     %f     %f\n", (double)p.XSHIFT, (double)p.YSHIFT,
            (double)p.HSCALE, (double)p.RKJDKI, (double)p.ANGORO);
    fputs("C8 INTERPOLATION SWITCHES\nC8 ISIRKI JSIRKI ISIHIHJ JSIHIHJ\n", g);
    fprintf(g, "   %d     %d     %d      %d\n", p.ISIRKI, p.JSIRKI, p.ISIHIHJ, p.JSIHIHJ);
    fputs("C9 NTYPE = 7 SPECIFID INPUT\nC9 IB IE JB JE N7RLX NXYIT ITN7M IJSMD ISMD JSMD RP7 SERRMAX\n", g);
    fputs("C10 NTYPE = 7 SPECIFID INPUT\nC10 X Y IN ORDER (IB,JB) (IE,JB) (IE,JE) (IB,JE)\n", g);
    fputs("C11 DEPTH INTERPOLATION SWITCHES\n"
          "C11 ISIDEP NDEPDAT CDEP RADM ISIDPTYP SURFELV ISVEG NVEGDAT NVEGTYP\n", g);
    fprintf(g, "    %d     %d      %f   %f   %d       %f      %d    %d      %d     %lf     %d\n",
            p.ISIDEP, ndepdat, (double)p.CDEP, (double)p.RADM, p.ISIDPTYP, (double)p.SURFELV, p.ISVEG,
            p.NVEGDAT, p.NVEGTYP, p.ZROUGH, p.VERBOSE);
    fputs("C12 LAST BOUNDARY POINT INFORMATION\nC12 ILT JLT X(ILT,JLT) Y(ILT,JLT)\n", g);
    fprintf(g, "    %d  %d       %f         %f\n", 1, 1, 0.0, 0.0);
    fputs("C13 BOUNDARY POINT INFORMATION\nC13 I J X(I,J) Y(I,J)\n", g);
    const bool ok = cell.close() && grid.close() && dep.close() && gef.close();
    if (!ok && err) *err = "cannot write the grid-generator inputs in " + dir;
    if (ndepdat_out) *ndepdat_out = ndepdat;
    return ok;
}

}  // namespace sw2dio
