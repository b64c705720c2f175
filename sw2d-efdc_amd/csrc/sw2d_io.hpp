// sw2d_io.hpp -- the reference's case-file contract (db/<case>/...), read and
// written by the sw2d-efdc driver.  Input parsing keeps the reference's
// fscanf semantics (whitespace in a format matches any run of whitespace,
// other characters must match literally, a failed conversion leaves the
// target unchanged); outputs keep its printf formats byte for byte.
#pragma once
#include <stdio.h>

#include <array>
#include <functional>
#include <string>
#include <vector>

namespace sw2dio {

// 1-SETUP and 2-SETUP blocks of setup.dat (SW2D-EFDC.cu:3763-3969)
struct Setup {
    double tday0 = 0, thour0 = 0, tmin0 = 0, tsec0 = 0;
    double tday = 0, thour = 0, tmin = 0, tsec = 0;
    double dt = 0, dpout = 0, dkout = 0, dtoq = 0;
    int evaporation_on = 0;
    double EV_WL_min = 0, INT = 0, INF = 0, LWL = 0, manning = 0;
    int out_vx = 0, out_vy = 0, out_elev = 0, out_depth = 0, out_outlet = 0;
    int efdc = 0;
    // EFDC block: only the values the SW2D side uses are kept by name
    double efdc_basetime = 0, TBEGIN = 0;
    int NTSPTC = 0, nlayers = 1;
    std::vector<double> efdc_block;  // every value of the block, in file order
};

// positions of named values in Setup::efdc_block: the order in which the
// reference reads setup.dat's EFDC block (SW2D-EFDC.cu:3823-3972)
enum EfdcValue {
    kBaseTime = 0, kTBEGIN, kNTSPTC, kNLayers, kZBRADJ, kZBRCVRT, kHMIN, kHADJ, kHCVRT, kHDRY,
    kHWET, kBELADJ, kBELCVRT, kNWSER, kNASER, kIASWRAD, kREVC, kRCHC, kSWRATNF, kSWRATNS,
    kFSWRATF, kDABEDT, kTBEDIT, kHTBED1, kHTBED2, kISPPH, kNPPPH, kISRPPH, kIPPHXY, kISVPH,
    kNPVPH, kISRVPH, kIVPHXY, kNTSSSS, kMTSP, kMTSC, kMTSA, kMTSUE, kMTSUT, kMTSU, kNSTYPE,
    kNSHOWR, kISHOWC, kJSHOWC, kNSHFREQ, kZSSMIN, kZSSMAX, kSSALMAX, kNProcessors, kLorpType,
    kDyeOn, kTempOn, kSaltOn, kDyeInit, kSaltInit, kTempBot, kTempSurf, kDyeIn, kSaltIn,
    kEfdcValues
};

struct Grid {
    int cols = 0, rows = 0;
    double xll = 0, yll = 0, cellsize = 0, nodata = 0;
    std::vector<double> v;  // rows*cols values, north row first
};

struct Case {
    std::string dir;  // db/<case>
    Setup setup;
    Grid dem;
    std::vector<double> src_x, src_y;  // coord_source_sink.dat (metres)
    std::vector<double> Q;             // Q_source_sink.dat, nq rows of nst values
    int nq = 0;
    double dtrain = 0;
    std::vector<double> rain;          // rain.dat, mm per dtrain (cont_rain - 1 values)
    std::vector<double> T, Rg;         // temperatures / solar radiation
    double albedo = 0;
    std::vector<double> h0;            // initial_condition.dat
    std::vector<double> out_x, out_y;  // coord_out.dat (metres)
    std::vector<int> mask;             // mask.dat (efdc == 1)
    Grid mask_hdr;                     // mask.dat header (efdc == 1)
    std::vector<double> patm, relh;    // atm_press.dat, rel_humidity.dat (efdc == 1)
    std::vector<double> wind;          // wind_velocities.dat pairs (efdc == 1, NWSER == 1)
    // efdc_coord_out.dat (efdc == 1, 4310-4334): EFDC time-series locations
    double efdc_tctmsr = 0;            // time_conversion (C84 TCTMSR)
    std::vector<double> efdc_out_x, efdc_out_y;  // metres
};

// count of '\n' in a file (the reference sizes series this way, 4041-4048)
int count_newlines(const std::string &path);

// db/dir.txt: case names
std::string read_dir_list(const std::string &root, std::vector<std::string> &names);

// read a whole case; returns "" or an error naming the missing file
std::string read_case(const std::string &root, const std::string &name, Case &c);

// VTK STRUCTURED_POINTS writer (SW2D-EFDC.cu:4958-5072): fields are rows*cols
// in memory order, written south row first with "%f"; a NULL field is skipped
struct VtkFields {
    const double *depth = nullptr, *vv = nullptr, *uu = nullptr, *z = nullptr;
};
bool write_vtk(const std::string &path, int cols, int rows, double dx, double dy, double x0,
               double y0, const VtkFields &f);

// one WaterLevel.txt output (4986-5013): probes are matched in memory order
// and print h at their cell with " %lf", the last one with a newline
void write_water_level(FILE *fp, int rows, int cols, double xcoor_cells, double ycoor_cells,
                       const std::vector<double> &outx_cells,
                       const std::vector<double> &outy_cells, const double *h);

// QSER.INP (write_efdc_QSER, 2499-2610) for outputs 0..tq; per edge list
// (0 N, 1 S, 2 E, 3 W): time[t] and val[t][record]
struct QserSeries {
    std::vector<double> time[4];
    std::vector<std::vector<double>> val[4];
    int nrec[4] = {0, 0, 0, 0};
};
bool write_qser(const std::string &path, int type, int nlayers, const QserSeries &q,
                double efdc_basetime, int tq);

// DSER.INP / TSER.INP / SSER.INP ('D', 'T', 'S'; write_efdc_DSER 2383-2498,
// TSER 2141-2254, SSER 2026-2140): one series per coupling record of q, in
// QSER's record order, value(t) in every layer at outputs 0..tq.  Every
// series takes its times from the north edge's list (the reference prints
// north_val[t][0] for all four edges).
bool write_constituent_series(const std::string &path, char kind, int type, int nlayers,
                              const QserSeries &q, const std::function<double(int)> &value,
                              double efdc_basetime, int tq);

// DYE.INP / SALT.INP / TEMP.INP (2255-2382): one line per lake cell
// (mask == 1), south row first, numbered in that order; value(cell, layer)
// with cell the memory index (north row first).
bool write_initial_field(const std::string &path, int cols, int rows, int nlayers, int water_cells,
                         const std::vector<int> &mask,
                         const std::function<double(size_t, int)> &value);

// WSER.INP (write_efdc_WSER 2612-2662): ntime (speed, direction) rows of wind
bool write_wser(const std::string &path, const std::vector<double> &wind, double data_basetime,
                double efdc_basetime, int ntime);

// ASER.INP (write_efdc_ASER 2664-2743): rows of 7 values (PATM, TDRY, RELH,
// RAIN, EVAP, SOLSWR, CLOUD); the ten thermal parameters in block order
// from kIASWRAD
bool write_aser(const std::string &path, const std::vector<double> &rows7, const double *thermal,
                double data_basetime, double efdc_basetime, int ntime);

// SHOW.INP (write_efdc_SHOW 2745-2764): the eight values from kNSTYPE
bool write_show(const std::string &path, const double *show);

// The values of the EFDC.INP cards that the coupled run rewrites
// (write_efdc_CONFIG, 1069-2024; arguments as passed at 4925-4933).
struct EfdcConfig {
    int nqsij = 0, nqser = 0;                 // coupling records (C23)
    std::vector<std::array<int, 2>> bc[4];    // (I, J) of the records, edges N, S, E, W (C24)
    int nlayers = 1, ncols = 0, nrows = 0, nwcells = 0;  // C9, C10
    double TCON = 0, TBEGIN = 0, TREF = 0;    // C8
    int NTC = 0, NTSPTC = 0, NTCPP = 0, NTCVB = 0;  // C7
    int ISPPH = 0, NPPPH = 0, ISRPPH = 0, IPPHXY = 0;  // C72
    int ISVPH = 0, NPVPH = 0, ISRVPH = 0, IVPHXY = 0;  // C73
    int ISTMSR = 0, MLTMSR = 0, NBTMSR = 0, NSTMSR = 0, NWTMSR = 0, NTSSTSP = 0;  // C84
    double TCTMSR = 0;
    std::vector<std::array<int, 2>> LTS;      // (I, J) of the time-series locations (C87)
    int NTSSSS = 0, MTSP = 0, MTSC = 0, MTSA = 0, MTSUE = 0, MTSUT = 0, MTSU = 0;
    double ZBRADJ = 0, ZBRCVRT = 0, HMIN = 0, HADJ = 0, HCVRT = 0, HDRY = 0, HWET = 0,
           BELADJ = 0, BELCVRT = 0;           // C11
    int NWSER = 0, NASER = 0;                 // C14
    double DX = 0, DY = 0;
    int dye_on = 0, salt_on = 0, temp_on = 0, temp_opt = 0;  // C6, C22, C24, C71
};

// Rewrites cards C6-C11, C14, C22-C26, C71-C73, C84 and C87 of the EFDC.INP
// at `path` in place of the template's; every byte outside those cards is
// kept.  false (and *err) when the file cannot be read or written or lacks
// one of the cards or their successors.
bool write_efdc_config(const std::string &path, const EfdcConfig &c, std::string *err = nullptr);

// grid_make_input.dat of a coupled case (the grid generator's controls,
// read at 2944-2991)
struct GefdcParams {
    int ITRXM = 0, ITRHM = 0, ITRKM = 0, ITRGM = 0, NDEPSM = 0, NDEPSMF = 0, DDATADJ = 0;
    float DEPMIN = 0, RPX = 0, RPK = 0, RPH = 0;
    double RSQXM = 0, RSQKM = 0, RSQKIM = 0, RSQHM = 0, RSQHIM = 0, RSQHJM = 0;
    float XSHIFT = 0, YSHIFT = 0, HSCALE = 0, RKJDKI = 0, ANGORO = 0;
    int ISIRKI = 0, JSIRKI = 0, ISIHIHJ = 0, JSIHIHJ = 0, ISIDEP = 0;
    float CDEP = 0, RADM = 0;
    int ISIDPTYP = 0;
    float SURFELV = 0;
    int ISVEG = 0, NVEGDAT = 0, NVEGTYP = 0;
    double ZROUGH = 0;
    int VERBOSE = 0;
};
bool read_grid_make_input(const std::string &path, GefdcParams &p, std::string *err = nullptr);

// CELL.INP cell types (0, 5 lake, 9 boundary) of a rows x cols mask, rows
// from the north (make_gefdc_inputs, 2877-3196)
std::vector<unsigned char> gefdc_cell_types(int cols, int rows, const std::vector<int> &mask);

// CELL.INP, gridext.inp, depdat.inp and gefdc.inp into directory `dir`
// (ending in '/') from the mask and the bed (rows x cols, north row first,
// after lake_catchment_smooth); *ndepdat receives the lake-cell count
bool write_gefdc_inputs(const std::string &dir, int cols, int rows, double res,
                        const std::vector<int> &mask, const std::vector<double> &bed,
                        const GefdcParams &p, int *ndepdat = nullptr, std::string *err = nullptr);

// EFDC's domain decomposition file (LORP.INP) for `nproc` ranks: the
// partition the reference's Gorp tool finds on these cell types (rows x
// cols, north row first, as gefdc_cell_types gives them) and the file it
// writes, for LORP type 1 (h: nproc columns of tiles), 2 (v: nproc rows) or
// 3 (r: the best grid of tiles); called at 4480-4507 of SW2D-EFDC.cu
bool write_lorp(const std::string &path, const std::vector<unsigned char> &types, int cols, int rows,
                int nproc, int kind, std::string *err = nullptr);

// cell types of a CELL.INP file, rows from the north, read the way the
// decomposition tool reads them (Gorp.cpp:17-60)
bool read_cell_types(const std::string &path, std::vector<unsigned char> &types, int &cols, int &rows,
                     std::string *err = nullptr);

}  // namespace sw2dio
