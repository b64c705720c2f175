// sw2d_io.cpp -- readers and writers of the reference's case files.
#include "sw2d_io.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>

namespace sw2dio {

namespace {

struct File {
    FILE *fp = nullptr;
    explicit File(const std::string &p, const char *mode) { fp = fopen(p.c_str(), mode); }
    ~File() { if (fp) fclose(fp); }
    explicit operator bool() const { return fp != nullptr; }
};

// One line of the setup.dat schema: a literal (the comment line, matched
// with fscanf semantics) optionally followed by one value.  The comment
// texts are the file format of setup.dat (SW2D-EFDC.cu:3763-3969).
enum Kind { LIT = 0, DBL = 1, INT = 2 };
struct Item {
    const char *lit;
    Kind kind;
};

const Item kSetup1[] = {
    {" #*******************************************************#", LIT},
    {" # 1-SETUP: SW2D-GPU                                     #", LIT},
    {" #*******************************************************#", LIT},
    {" # START TIME: DAY, HOUR, MINUTES, SECONDS #", LIT},  // 4 values follow
};

const Item kSetup1Tail[] = {
    {" # TIME STEP #", DBL},
    {" # TIME INTERVAL TO SHOW TIME ON SCREEM (SECONDS) #", DBL},
    {" # TIME INTERVAL TO WRITE THE RESULTS (SECONDS) #", DBL},
    {" # TEMPORAL RESOLUTION OF FLOW INPUT DATA (SECONDS) #", DBL},
    {" # EVAPORATION IN THE LAKE: 1 ACTIVE, 0 INACTIVE #", INT},
    {" # MINIMUM WATER DEPTH FOR EVAPORATION #", DBL},
    {" # PERCENTAGE OF WATER LOST BY INTERCEPTION #", DBL},
    {" # PERCENTAGE OF WATER LOST BY INFILTRATION #", DBL},
    {" # PERCENTAGE OF WATER LOSS IN THE LAKE #", DBL},
    {" # MANNING COEFFICIENT #", DBL},
    {" # OUTPUT X VELOCITY: 1 ACTIVE, 0 INACTIVE #", INT},
    {" # OUTPUT Y VELOCITY: 1 ACTIVE, 0 INACTIVE #", INT},
    {" # OUTPUT ELEVATIONS: 1 ACTIVE, 0 INACTIVE #", INT},
    {" # OUTPUT WATER DEPTH: 1 ACTIVE, 0 INACTIVE #", INT},
    {" # OUTPUT WATER DEPTH TIME SERIES: 1 ACTIVE, 0 INACTIVE #", INT},
    {" #*******************************************************#", LIT},
    {" # 2-SETUP: EFDC-MPI                                     #", LIT},
    {" #*******************************************************#", LIT},
    {" # 3D HYDRODYNAMIC MODEL: 1 ACTIVE, 0 INACTIVE #", INT},
};

// the EFDC block (read only when the 3D model flag is 1)
const Item kSetupEfdc[] = {
    {" #-------------- TEMPORAL CONTROLLERS -------------------#", LIT},
    {" # EFDC BASE TIME #", DBL},
    {" # TIME ORIGIN OF RUN #", DBL},
    {" # NUMBER OF TIME STEPS PER REFERENCE TIME PERIOD #", INT},
    {" #------- ROUGHNESS AND DEPTH PARAMETERS ----------#", LIT},
    {" # NUMBER OF VERTICAL LAYERS #", INT},
    {" # LOG BDRY LAYER CONST OR VARIABLE ROUGH HEIGHT ADJ IN METERS #", DBL},
    {" # LOG BDRY LAYER VARIABLE ROUGHNESS HEIGHT CONVERT TO METERS #", DBL},
    {" # MINIMUM DEPTH OF INPUTS DEPTHS IN METERS #", DBL},
    {" # ADJUCTMENT TO DEPTH FIELD IN METERS #", DBL},
    {" # CONVERTS INPUT DEPTH FIELD TO METERS #", DBL},
    {" # DEPTH AT WHICH CELL OR FLOW FACE BECOMES DRY #", DBL},
    {" # DEPTH AT WHICH CELL OR FLOW FACE BECOMES WET #", DBL},
    {" # ADJUCTMENT TO BOTTOM BED ELEVATION FIELD IN METERS #", DBL},
    {" # CONVERTS INPUT BOTTOM BED ELEVATION FIELD TO METERS #", DBL},
    {" #----------------- ATMOSPHERIC FORCING -----------------#", LIT},
    {" # NUMBER OF WIND TIME SERIES (0 SETS WIND TO ZERO) #", INT},
    {" # NUMBER OF ATMOSPHERIC CONDITION TIME SERIES (0 SETS ALL  ZERO) #", INT},
    {" #---------- ATMOSPHERIC AND THERMAL PROPERTIES ---------#", LIT},
    {" # 0 DISTRIBUTE SW SOL RAD OVER WATER COL AND INTO BED, = 1 ALL TO SURF LAYER #", INT},
    {" # 1000 * EVAPORATIVE TRANSFER COEF, REVC<0 USE WIND SPD DEPD DRAG COEF #", DBL},
    {" # 1000 * CONVECTIVE HEAT TRANSFER COEF, REVC<0 USE WIND SPD DEPD DRAG COEF #", DBL},
    {" # FAST SCALE SOLAR SW RADIATION ATTENUATION COEFFCIENT 1./METERS #", DBL},
    {" # SLOW SCALE SOLAR SW RADIATION ATTENUATION COEFFCIENT 1./METERS #", DBL},
    {" # FRACTION OF SOLSR SW RADIATION ATTENUATED FAST  0<FSWRATF<1 #", DBL},
    {" # DEPTH OR THICKNESS OF ACTIVE BED TEMPERATURE LAYER, METERS #", DBL},
    {" # INITIAL BED TEMPERATURE #", DBL},
    {" # CONVECTIVE HT COEFFCIENT BETWEEN BED AND BOTTOM WATER LAYER  NO DIM #", DBL},
    {" # HEAT TRANS COEFFCIENT BETWEEN BED AND BOTTOM WATER LAYER  M/SEC #", DBL},
    {" #----- CONTROLS FOR HORIZONTAL SURFACE ELEVATION OR PRESSURE CONTOURING -----#", LIT},
    {" # SURFACE ELEVATION OUTPUT #", LIT},
    {" # 1 TO WRITE FILE FOR SURFACE ELEVATION OR PRESSURE CONTOURING #", LIT},
    {" # 2 WRITE ONLY DURING LAST REFERENCE TIME PERIOD #", INT},
    {" # NUMBER OF WRITES PER REFERENCE TIME PERIOD #", INT},
    {" # 1 TO WRITE FILE FOR RESIDUAL SURFACE ELEVATION  CONTOURNG IN HORIZONTAL PLANE #", INT},
    {" # SURFACE ELEVATION WRITING FORMAT #", LIT},
    {" # 0 DOES NOT WRITE I,J,X,Y IN surfplt.out and rsurfplt.out FILES #", LIT},
    {" # 1 WRITES I,J ONLY IN surfplt.out and rsurfplt.out FILES #", LIT},
    {" # 2 WRITES I,J,X,Y  IN surfplt.out and rsurfplt.out FILES #", LIT},
    {" # 3 WRITES EFDC EXPLORER BINARY FORMAT FILES #", INT},
    {" #------ CONTROLS FOR HORIZONTAL PLANE VELOCITY VECTOR PLOTTING ----------#", LIT},
    {" # HORIZONTAL PLANE VELOCITY OUTPUT #", LIT},
    {" # 1 TO WRITE FILE FOR VELOCITY PLOTTING IN HORIZONTAL PLANE #", LIT},
    {" # 2 WRITE ONLY DURING LAST REFERENCE TIME PERIOD #", INT},
    {" # NUMBER OF WRITES PER REFERENCE TIME PERIOD #", INT},
    {" # 1 TO WRITE FILE FOR RESIDUAL VELOCITY PLOTTIN IN HORIZONTAL PLANE #", INT},
    {" # HORIZONTAL VELOCITY WRITING FORMAT #", LIT},
    {" # 0 DOES NOT WRITE I,J,X,Y IN velplth.out and rvelplth.out FILES #", LIT},
    {" # 1 WRITES I,J ONLY IN velplth.out and rvelplth.out FILES #", LIT},
    {" # 2 WRITES I,J,X,Y  IN velplth.out and rvelplth.out FILES #", LIT},
    {" # 3 WRITES EFDC EXPLORER BINARY FORMAT FILES #", INT},
    {" #--------------- CONTROLS FOR WRITING TO TIME SERIES FILES --------------#", LIT},
    {" # WRITE SCENARIO FOR THIS LOCATION #", INT},
    {" # 1 FOR TIME SERIES OF SURFACE ELEVATION #", INT},
    {" # 1 FOR TIME SERIES OF TRANSPORTED CONCENTRATION VARIABLES #", INT},
    {" # 1 FOR TIME SERIES OF EDDY VISCOSITY AND DIFFUSIVITY #", INT},
    {" # 1 FOR TIME SERIES OF EXTERNAL MODE HORIZONTAL VELOCITY #", INT},
    {" # 1 FOR TIME SERIES OF EXTERNAL MODE HORIZONTAL TRANSPORT #", INT},
    {" # 1 FOR TIME SERIES OF HORIZONTAL VELOCITY IN EVERY LAYER #", INT},
    {" #---------------------- SHOW.INP PARAMETERS ---------------------------#", LIT},
    {" # NSTYPE #", INT},
    {" # NSHOWR #", INT},
    {" # ISHOWC #", INT},
    {" # JSHOWC #", INT},
    {" # NSHFREQ #", INT},
    {" # ZSSMIN #", INT},
    {" # ZSSMAX #", INT},
    {" # SSALMAX #", INT},
    {" #----------------- DOMAIN DECOMPOSITION PARAMETERS --------------------#", LIT},
    {" # NUMBER OF PROCESSORS #", INT},
    {" # LOCALLY OPTIMAL RECTILINEAR PARTITION (LORP) TYPE #", LIT},
    {" # 1 TYPE h #", LIT},
    {" # 2 TYPE v #", LIT},
    {" # 3 TYPE r #", INT},
    {" #---------------- CHOOSE TO INCLUDE IN THE SIMULATION -----------------#", LIT},
    {" # DYE SIMULATION: 1 ACTIVE, 0 INACTIVE #", INT},
    {" # TEMPERATURE SIMULATION: 1 ACTIVE, 0 INACTIVE #", INT},
    {" # SALINITY SIMULATION: 1 ACTIVE, 0 INACTIVE #", INT},
    {" #----------------------- INITIAL CONDITIONS ---------------------------#", LIT},
    {" # INITIAL DYE CONCENTRATION IN THE LAKE #", DBL},
    {" # INITIAL SALINITY IN THE LAKE #", DBL},
    {" # WATER TEMPERATURE AT THE BOTTOM OF THE LAKE #", DBL},
    {" # WATER TEMPERATURE ON THE LAKE SURFACE #", DBL},
    {" #---- CATCHMENT-LAKE BOUNDARIES (WATER INPUTS FROM THE CATCHMENT) -----#", LIT},
    {" # DYE CONCENTRATION #", DBL},
    {" # SALINITY #", DBL},
};

}  // namespace

// fscanf with a literal-only format: whitespace matches any run of
// whitespace, every other character must match; stops at the first mismatch
static void scan_lit(FILE *fp, const char *lit)
{
    if (strchr(lit, '%')) return;  // the schema holds no conversions here
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wformat-security"
    int r = fscanf(fp, lit);
#pragma GCC diagnostic pop
    (void)r;
}

static void scan_value(FILE *fp, Kind kind, double *out)
{
    if (kind == DBL) {
        double v = *out;
        if (fscanf(fp, " %lf\n ", &v) == 1) *out = v;
    } else if (kind == INT) {
        int v = (int)*out;
        if (fscanf(fp, " %d\n ", &v) == 1) *out = v;
    }
}

static void read_setup(FILE *fp, Setup &s)
{
    for (const Item &it : kSetup1) scan_lit(fp, it.lit);
    if (fscanf(fp, " %lf %lf %lf %lf\n ", &s.tday0, &s.thour0, &s.tmin0, &s.tsec0) < 4) {}
    scan_lit(fp, " # END TIME: DAY, HOUR, MINUTES, SECONDS #");
    if (fscanf(fp, " %lf %lf %lf %lf\n ", &s.tday, &s.thour, &s.tmin, &s.tsec) < 4) {}
    double v[sizeof(kSetup1Tail) / sizeof(kSetup1Tail[0])];
    int n = 0;
    for (const Item &it : kSetup1Tail) {
        scan_lit(fp, it.lit);
        v[n] = 0;
        scan_value(fp, it.kind, &v[n]);
        n++;
    }
    // value order of kSetup1Tail (LIT entries hold a 0)
    s.dt = v[0]; s.dpout = v[1]; s.dkout = v[2]; s.dtoq = v[3];
    s.evaporation_on = (int)v[4]; s.EV_WL_min = v[5]; s.INT = v[6]; s.INF = v[7];
    s.LWL = v[8]; s.manning = v[9];
    s.out_vx = (int)v[10]; s.out_vy = (int)v[11]; s.out_elev = (int)v[12];
    s.out_depth = (int)v[13]; s.out_outlet = (int)v[14];
    s.efdc = (int)v[18];
    if (s.efdc == 1) {
        s.efdc_block.clear();
        for (const Item &it : kSetupEfdc) {
            scan_lit(fp, it.lit);
            if (it.kind != LIT) {
                double x = 0;
                scan_value(fp, it.kind, &x);
                s.efdc_block.push_back(x);
            }
        }
        s.efdc_basetime = s.efdc_block[0];
        s.TBEGIN = s.efdc_block[1];
        s.NTSPTC = (int)s.efdc_block[2];
        s.nlayers = (int)s.efdc_block[3];
    }
}

static bool read_esri_header(FILE *fp, Grid &g)
{
    int ok = 0;
    ok += fscanf(fp, " ncols %d\n", &g.cols);
    ok += fscanf(fp, " nrows %d\n", &g.rows);
    ok += fscanf(fp, " xllcorner %lf\n", &g.xll);
    ok += fscanf(fp, " yllcorner %lf\n", &g.yll);
    ok += fscanf(fp, " cellsize %lf\n", &g.cellsize);
    ok += fscanf(fp, " NODATA_value %lf\n", &g.nodata);
    return ok == 6 && g.cols > 0 && g.rows > 0;
}

int count_newlines(const std::string &path)
{
    File f(path, "r");
    if (!f) return -1;
    int n = 0, ch;
    while ((ch = fgetc(f.fp)) != EOF)
        if (ch == '\n') n++;
    return n;
}

std::string read_dir_list(const std::string &root, std::vector<std::string> &names)
{
    File f(root + "/db/dir.txt", "r");
    if (!f) return "unknown file - db/dir.txt";
    int n = 0;
    if (fscanf(f.fp, " dir_number: %d", &n) != 1) return "malformed db/dir.txt";
    char buf[4000];
    for (int i = 0; i < n; i++) {
        if (fscanf(f.fp, " %3999s", buf) != 1) return "db/dir.txt lists fewer cases than dir_number";
        names.push_back(buf);
    }
    return "";
}

std::string read_case(const std::string &root, const std::string &name, Case &c)
{
    c.dir = root + "/db/" + name;
    const std::string in = c.dir + "/SW2D/input/";
    {
        File f(c.dir + "/setup.dat", "r");
        if (!f) return "unknown file - setup.dat";
        read_setup(f.fp, c.setup);
    }
    {
        File f(in + "dem.dat", "r");
        if (!f) return "unknown file - dem.dat";
        if (!read_esri_header(f.fp, c.dem)) return "malformed dem.dat header";
        const size_t n = (size_t)c.dem.rows * c.dem.cols;
        c.dem.v.assign(n, 0.0);
        for (size_t i = 0; i < n; i++)
            if (fscanf(f.fp, "%lf\n", &c.dem.v[i]) != 1) return "dem.dat holds fewer than nrows*ncols values";
    }
    int nst = 0;
    {
        File f(in + "coord_source_sink.dat", "r");
        if (!f) return "unknown file - coord_source_sink.dat";
        if (fscanf(f.fp, "nst  %d\n ", &nst) != 1 || nst < 0) return "malformed coord_source_sink.dat";
        for (int i = 0; i < nst; i++) {
            double x = 0, y = 0;
            if (fscanf(f.fp, " %lf %lf\n ", &x, &y) != 2) return "coord_source_sink.dat holds fewer than nst points";
            c.src_x.push_back(x);
            c.src_y.push_back(y);
        }
    }
    {
        const std::string p = in + "Q_source_sink.dat";
        const int cont_qq = count_newlines(p);
        if (cont_qq < 0) return "unknown file - Q_source_sink.dat";
        File f(p, "r");
        char skip[4096];
        if (fscanf(f.fp, "flow %4095s\n", skip) != 1) {}
        c.Q.clear();
        c.nq = 0;
        // the reference reads cont_qq rows; rows past the end of the file stay 0
        for (int r = 0; r < cont_qq; r++) {
            bool any = false;
            for (int i = 0; i < nst; i++) {
                double v = 0;
                if (fscanf(f.fp, " %lf\n", &v) == 1) any = true;
                c.Q.push_back(v);
            }
            if (any || nst == 0) c.nq = r + 1;
        }
        c.Q.resize((size_t)c.nq * nst);
    }
    {
        const std::string p = in + "rain.dat";
        const int cont_rain = count_newlines(p);
        if (cont_rain < 0) return "unknown file - rain.dat";
        File f(p, "r");
        if (fscanf(f.fp, " %lf\n", &c.dtrain) != 1) return "malformed rain.dat";
        char h1[256], h2[256], h3[256], h4[256];
        if (fscanf(f.fp, " %255s %255s %255s %255s\n", h1, h2, h3, h4) != 4) return "malformed rain.dat header";
        c.rain.assign(cont_rain > 1 ? cont_rain - 1 : 0, 0.0);
        char label[256];
        for (size_t i = 0; i < c.rain.size(); i++) {
            double v = 0;
            if (fscanf(f.fp, " %255s %lf\n", label, &v) == 2) c.rain[i] = v;
        }
        if (c.setup.evaporation_on == 1) {
            File ft(in + "temperatures.dat", "r");
            File fs(in + "solar_radiation.dat", "r");
            if (!ft || !fs) return "unknown file - evaporation.dat";
            char head[256];
            if (fscanf(fs.fp, " albedo: %lf\n", &c.albedo) != 1) return "malformed solar_radiation.dat";
            if (fscanf(fs.fp, " %255s\n", head) != 1) {}
            if (fscanf(ft.fp, " %255s\n", head) != 1) {}
            c.T.assign(c.rain.size(), 0.0);
            c.Rg.assign(c.rain.size(), 0.0);
            for (size_t i = 0; i < c.rain.size(); i++) {
                if (fscanf(ft.fp, " %lf\n", &c.T[i]) != 1) {}
                if (fscanf(fs.fp, " %lf\n", &c.Rg[i]) != 1) {}
            }
        }
    }
    {
        File f(in + "initial_condition.dat", "r");
        if (!f) return "unknown file - initial_condition.txt";
        const size_t n = (size_t)c.dem.rows * c.dem.cols;
        c.h0.assign(n, 0.0);
        for (size_t i = 0; i < n; i++)
            if (fscanf(f.fp, "%lf\n", &c.h0[i]) != 1) return "initial_condition.dat holds fewer than nrows*ncols values";
    }
    if (c.setup.out_outlet == 1) {
        File f(in + "coord_out.dat", "r");
        if (!f) return "unknown file - coord_out.dat";
        int n = 0;
        if (fscanf(f.fp, "n_out  %d\n ", &n) != 1 || n < 0) return "malformed coord_out.dat";
        for (int i = 0; i < n; i++) {
            double x = 0, y = 0;
            if (fscanf(f.fp, " %lf %lf\n ", &x, &y) != 2) return "coord_out.dat holds fewer than n_out points";
            c.out_x.push_back(x);
            c.out_y.push_back(y);
        }
    }
    if (c.setup.efdc == 1) {
        File f(in + "mask.dat", "r");
        if (!f) return "unknown file - mask.dat";
        if (!read_esri_header(f.fp, c.mask_hdr)) return "malformed mask.dat header";
        if (c.mask_hdr.rows != c.dem.rows || c.mask_hdr.cols != c.dem.cols)
            return "mask.dat and dem.dat differ in size";
        const size_t n = (size_t)c.dem.rows * c.dem.cols;
        c.mask.assign(n, 0);
        for (size_t i = 0; i < n; i++)
            if (fscanf(f.fp, "%d\n", &c.mask[i]) != 1) return "mask.dat holds fewer than nrows*ncols values";
    }
    return "";
}

bool write_vtk(const std::string &path, int cols, int rows, double dx, double dy, double x0,
               double y0, const VtkFields &f)
{
    File o(path, "w");
    if (!o) return false;
    FILE *fp = o.fp;
    fprintf(fp, "%s\n", "# vtk DataFile Version 2.0");
    fprintf(fp, "%s\n", "Brazil");
    fprintf(fp, "%s\n", "ASCII");
    fprintf(fp, "%s\n", "DATASET STRUCTURED_POINTS");
    fprintf(fp, "DIMENSIONS %d %d %d\n", cols, rows, 1);
    fprintf(fp, "ASPECT_RATIO %lf %lf %lf\n", dx, dy, 1.0000);
    fprintf(fp, "ORIGIN %lf %lf %lf\n", x0, y0, 0.000);
    fprintf(fp, "POINT_DATA %d\n", cols * rows);
    struct Block { const char *name; const double *v; double sign; };
    const Block blocks[] = {{"depth", f.depth, 1.0}, {"x_velocity", f.vv, 1.0},
                            {"y_velocity", f.uu, -1.0}, {"Elevations", f.z, 1.0}};
    for (const Block &b : blocks) {
        if (!b.v) continue;
        fprintf(fp, "SCALARS %s float 1\n", b.name);
        fprintf(fp, "%s\n", "LOOKUP_TABLE default");
        for (int im = 0; im < rows; im++) {
            const double *row = b.v + (size_t)(rows - 1 - im) * cols;
            for (int jm = 0; jm < cols; jm++)
                fprintf(fp, "%f\n", b.sign < 0 ? -row[jm] : row[jm]);
        }
    }
    return true;
}

void write_water_level(FILE *fp, int rows, int cols, double xcoor_cells, double ycoor_cells,
                       const std::vector<double> &outx_cells,
                       const std::vector<double> &outy_cells, const double *h)
{
    // probe i sits at flat index (outy - 1) * cols + outx + 1 (4989-4992);
    // the reference scans the cells in memory order and prints every probe
    // that matches, so probes come out ordered by cell
    const int n_out = (int)outx_cells.size();
    std::vector<long> idx(n_out);
    for (int i = 0; i < n_out; i++) {
        const long outx = lround(fabs(xcoor_cells - outx_cells[i]));
        const long outy = rows - lround(fabs(ycoor_cells - outy_cells[i]));
        idx[i] = outy * cols - (cols - (outx + 1));
    }
    std::vector<std::pair<long, int>> order;
    for (int i = 0; i < n_out; i++)
        if (idx[i] >= 0 && idx[i] < (long)rows * cols) order.emplace_back(idx[i], i);
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<long, int> &a, const std::pair<long, int> &b) { return a.first < b.first; });
    int npout = 0;
    for (const auto &o : order) {
        npout++;
        if (npout == n_out) fprintf(fp, " %lf\n", h[o.first]);
        else fprintf(fp, " %lf", h[o.first]);
    }
}

bool write_qser(const std::string &path, int type, int nlayers, const QserSeries &q,
                double efdc_basetime, int tq)
{
    File o(path, "w");
    if (!o) return false;
    FILE *fp = o.fp;
    static const char *hdr[] = {
        "C **, qser.inp Time Series FILE",
        "C **",
        "C **  InType MQSER(NS) TCQSER(NS) TAQSER(NS) RMULADJ(NS) ADDADJ(NS) ICHGQS",
        "C **",
        "C **  IF InType.EQ.1 THEN READ DEPTH WEIGHTS AND SINGLE VALUE OF QSER",
        "C **                      ELSE READ A VALUE OF QSER FOR EACH LAYER",
        "C **",
        "C **  InType = 1 Structure",
        "C **  WKQ(K), K = 1, KC",
        "C **  TQSER(M, NS)    QSER(M, 1, NS)          !(MQSER(NS) PAIRS FOR NS = 1, NQSER SERIES)",
        "C **",
        "C **  InType = 0 Structure",
        "C **  TQSER(M, NS)    (QSER(M, K, NS), K = 1, KC) !(MQSER(NS) PAIRS)",
        "C **",
    };
    for (const char *l : hdr) fprintf(fp, "%s\n", l);
    static const char *names[4] = {" ' *** North_edge_", " ' *** South_edge_", " ' *** East_edge_",
                                   " ' *** West_edge_"};
    int edg = 0;
    for (int e = 0; e < 4; e++) {
        for (int b = 0; b < q.nrec[e]; b++) {
            edg++;
            fprintf(fp, "      %d %d %lf %d %d %d %d %s%d\n", type, tq + 1, efdc_basetime, 0, 1, 0,
                    0, names[e], edg);
            for (int t = 0; t <= tq; t++) {
                fprintf(fp, "      %lf ", q.time[e][t]);
                const double v = q.val[e][t][b];
                for (int il = 0; il < nlayers; il++) {
                    if (il == nlayers - 1) fprintf(fp, "%lf\n", v / nlayers);
                    else fprintf(fp, "%lf ", v / nlayers);
                }
            }
        }
    }
    return true;
}

}  // namespace sw2dio
