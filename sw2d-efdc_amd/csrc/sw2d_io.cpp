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
        if (s.efdc_block.size() < (size_t)kEfdcValues) s.efdc_block.resize(kEfdcValues, 0.0);
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
        // EFDC time-series locations (4310-4334)
        File g(in + "efdc_coord_out.dat", "r");
        if (!g) return "unknown file - efdc_coord_out.dat";
        int n = 0;
        if (fscanf(g.fp, "time_conversion:  %lf\n ", &c.efdc_tctmsr) != 1 ||
            fscanf(g.fp, "number_t_s_out:  %d\n ", &n) != 1 || n < 0)
            return "malformed efdc_coord_out.dat";
        for (int i = 0; i < n; i++) {
            double x = 0, y = 0;
            if (fscanf(g.fp, " %lf %lf\n ", &x, &y) != 2)
                return "efdc_coord_out.dat holds fewer than number_t_s_out points";
            c.efdc_out_x.push_back(x);
            c.efdc_out_y.push_back(y);
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
        // EFDC meteorological series (4427-4455): a header word, then one
        // value per data interval; wind (4405-4425) as speed/direction pairs
        auto series = [&](const char *file, const char *head, std::vector<double> &out, int per) {
            File g(in + file, "r");
            if (!g) return false;
            if (head) scan_lit(g.fp, head);
            double v[2];
            while (fscanf(g.fp, per == 2 ? " %lf %lf\n " : " %lf\n ", &v[0], &v[1]) == per)
                out.insert(out.end(), v, v + per);
            return true;
        };
        if (!series("atm_press.dat", "atm_pressure(millibar)", c.patm, 1)) return "unknown file - atm_press.dat";
        if (!series("rel_humidity.dat", "rel_humidity(percent)", c.relh, 1)) return "unknown file - rel_humidity.dat";
        if ((int)c.setup.efdc_block[kNWSER] == 1 && !series("wind_velocities.dat", nullptr, c.wind, 2))
            return "unknown file - wind_velocities.dat";
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

namespace sw2dio {

bool write_constituent_series(const std::string &path, char kind, int type, int nlayers,
                              const QserSeries &q, const std::function<double(int)> &value,
                              double efdc_basetime, int tq)
{
    File o(path, "w");
    if (!o) return false;
    FILE *fp = o.fp;
    // the header of EFDC's constituent series files: the QSER layout with
    // the constituent's letter in the series names
    const char k = kind;
    fprintf(fp, "C **, %cSER.INP Time Series FILE\nC **\nC **\n", k);
    fprintf(fp, "C **  InType M%cSER(NS) TC%cSER(NS) TA%cSER(NS) RMULADJ(NS) ADDADJ(NS)\nC **\n", k, k, k);
    fprintf(fp, "C **  IF InType.EQ.1 THEN READ DEPTH WEIGHTS AND SINGLE VALUE OF %cSER\n", k);
    fprintf(fp, "C **                      ELSE READ A VALUE OF %cSER FOR EACH LAYER\nC **\n", k);
    fprintf(fp, "C **  InType = 1 Structure\nC **  WKQ(K), K = 1, KC\n");
    fprintf(fp, "C **  T%cSER(M, NS)    %cSER(M, 1, NS) !(M%cSER(NS) PAIRS FOR NS = 1, N%cSER SERIES)\nC **\n",
            k, k, k, k);
    fprintf(fp, "C **  InType = 0 Structure\n");
    fprintf(fp, "C **  T%cSER(M, NS)    (%cSER(M, K, NS), K = 1, KC) !(M%cSER(NS) PAIRS)\nC **\n", k, k, k);
    static const char *names[4] = {" ' *** North_edge_", " ' *** South_edge_", " ' *** East_edge_",
                                   " ' *** West_edge_"};
    int edg = 0;
    for (int e = 0; e < 4; e++) {
        for (int b = 0; b < q.nrec[e]; b++) {
            edg++;
            fprintf(fp, "      %d %d %lf %d %d %d %s%d\n", type, tq + 1, efdc_basetime, 0, 1, 0, names[e], edg);
            for (int t = 0; t <= tq; t++) {
                const double tt = t < (int)q.time[0].size() ? q.time[0][t] : 0.0;
                fprintf(fp, "      %lf ", tt);
                const double v = value(t);
                for (int il = 0; il < nlayers; il++) fprintf(fp, il == nlayers - 1 ? "%lf\n" : "%lf ", v);
            }
        }
    }
    return true;
}

bool write_initial_field(const std::string &path, int cols, int rows, int nlayers, int water_cells,
                         const std::vector<int> &mask,
                         const std::function<double(size_t, int)> &value)
{
    File o(path, "w");
    if (!o) return false;
    FILE *fp = o.fp;
    // the cell number's field width: the reference's expression of the digit
    // count of the lake's cell count, minus one
    const int width = water_cells > 0 ? (int)(round(1 + log((double)water_cells) / log(10.0)) - 1) : 0;
    fprintf(fp, "C\nC\nC\nC\n1\n");
    int cell = 0;
    for (int jr = 0; jr < rows; jr++) {
        const size_t row0 = (size_t)(rows - 1 - jr) * cols;   // south row first
        for (int i = 0; i < cols; i++) {
            const size_t m = row0 + i;
            if (mask[m] != 1) continue;
            cell++;
            fprintf(fp, "%*d     %d     %d     ", width, cell, i + 1, jr + 1);
            for (int l = 0; l < nlayers; l++) fprintf(fp, l < nlayers - 1 ? "%lf    " : "%lf\n", value(m, l));
        }
    }
    return true;
}

bool write_wser(const std::string &path, const std::vector<double> &wind, double data_basetime,
                double efdc_basetime, int ntime)
{
    if ((size_t)2 * std::max(ntime, 0) > wind.size()) return false;
    File o(path, "w");
    if (!o) return false;
    FILE *fp = o.fp;
    static const char *hdr[] = {
        "C **, wser.inp Time Series FILE",
        "C **",
        "C **  WIND FORCING FILE, USE WITH 7 APRIL 97 AND LATER VERSIONS OF EFDC",
        "C **",
        "C **  MASER(NW) = NUMBER OF TIME DATA POINTS",
        "C **  TCASER(NW) = DATA TIME UNIT CONVERSION TO SECONDS",
        "C **  TAASER(NW) = ADDITIVE ADJUSTMENT OF TIME VALUES SAME UNITS AS INPUT TIMES",
        "C **  WINDSCT(NW) = WIND SPEED CONVERSION TO M / SEC",
        "C **  ISWDINT(NW) = DIRECTION CONVENTION",
        "C **               0 DIRECTION TO",
        "C **               1 DIRECTION FROM",
        "C **               2 WINDS IS EAST VELOCITY, WINDD IS NORTH VELOCITY",
        "#  EFDC_DS_WIND_HEIGHT: 10.0",
        "#  EFDC_DS_ICE_PERIOD_START : 0.",
        "#  EFDC_DS_ICE_PERIOD_END : 0.",
        "C **  MASER  TCASER   TAASER  WINDSCT  ISWDINT",
        "C **  TASER(M) WINDS(M) WINDD(M)",
    };
    for (const char *l : hdr) fprintf(fp, "%s\n", l);
    fprintf(fp, "      %d %lf %lf %lf %d\n", ntime, efdc_basetime, 0.0, 1.0, 1);
    for (int t = 0; t < ntime; t++)
        fprintf(fp, "      %lf %lf %lf\n", ((double)t * data_basetime) / efdc_basetime, wind[2 * t], wind[2 * t + 1]);
    return true;
}

bool write_aser(const std::string &path, const std::vector<double> &rows7, const double *thermal,
                double data_basetime, double efdc_basetime, int ntime)
{
    if ((size_t)7 * std::max(ntime, 0) > rows7.size()) return false;
    File o(path, "w");
    if (!o) return false;
    FILE *fp = o.fp;
    static const char *hdr[] = {
        "# ASER.INP ",
        "# ",
        "#\t",
        "#  ATMOSPHERIC FORCING FILE, USE WITH 28 JULY 96 AND LATER VERSIONS OF EFDC",
        "#  MASER = NUMBER OF TIME DATA POINTS",
        "#  TCASER = DATA TIME UNIT CONVERSION TO SECONDS",
        "#  TAASER = ADDITIVE ADJUSTMENT OF TIME VALUES SAME UNITS AS INPUT TIMES",
        "#  IRELH = 0 VALUE TWET COLUMN VALUE IS TWET, = 1 VALUE IS RELATIVE HUMIDITY",
        "#  RAINCVT = CONVERTS RAIN TO UNITS OF M / SEC.in / hr * 0.0254 m / 1 in * 1 hr / 3600 s = 7.05556E-06 (m / s)",
        "#  EVAPCVT = CONVERTS EVAP TO UNITS OF M / SEC, IF EVAPCVT<0 EVAP IS INTERNALLY COMPUTED",
        "#  SOLRCVT = CONVERTS SOLAR SW RADIATION TO JOULES / SQ METER",
        "#  CLDCVT = MULTIPLIER FOR ADJUSTING CLOUD COVER",
        "#  IASWRAD = 0 DISTRIBUTE SW SOL RAD OVER WATER COL AND INTO BED, = 1 ALL TO SURF LAYER",
        "#  REVC = 1000 * EVAPORATIVE TRANSFER COEF, REVC<0 USE WIND SPD DEPD DRAG COEF",
        "#  RCHC = 1000 * CONVECTIVE HEAT TRANSFER COEF, REVC<0 USE WIND SPD DEPD DRAG COEF",
        "#  SWRATNF = FAST SCALE SOLAR SW RADIATION ATTENUATION COEFFCIENT 1. / METERS",
        "#  SWRATNS = SLOW SCALE SOLAR SW RADIATION ATTENUATION COEFFCIENT 1. / METERS",
        "#  FSWRATF = FRACTION OF SOLSR SW RADIATION ATTENUATED FAST  0<FSWRATF<1",
        "#  DABEDT = DEPTH OR THICKNESS OF ACTIVE BED TEMPERATURE LAYER, METERS",
        "#  TBEDIT = INITIAL BED TEMPERATURE",
        "#  HTBED1 = CONVECTIVE HT COEFFCIENT BETWEEN BED AND BOTTOM WATER LAYER  NO DIM",
        "#  HTBED2 = HEAT TRANS COEFFCIENT BETWEEN BED AND BOTTOM WATER LAYER  M / SEC",
        "#  PATM = ATM PRESS MILLIBAR",
        "#  TDRY / TEQ = DRY ATM TEMP ISOPT(2) = 1 OR EQUIL TEMP ISOPT(2) = 2",
        "#  TWET / RELH = WET BULB ATM TEMP IRELH = 0, RELATIVE HUMIDITY IRELH = 1",
        "#  RAIN = RAIN FALL RATE LENGTH / TIME",
        "#  EVAP = EVAPORATION RATE IF EVAPCVT>0.",
        "#  SOLSWR = SOLAR SHORT WAVE RADIATION AT WATER SURFACE  ENERGY FLUX / UNIT AREA",
        "#  CLOUD = FRATIONAL CLOUD COVER",
        "#",
        "#  MASER    TCASER  TAASER  IRELH    RAINCVT  EVAPCVT  SOLRCVT   CLDCVT",
        "#",
        "#  IASWRAD  REVC    RCHC    SWRATNF  SWRATNS  FSWRATF  DABEDT    TBEDIT    HTBED1    HTBED2",
        "#",
        "#  TASER(M) PATM(M) TDRY(M) TWET(M)  RAIN(M)  EVAP(M)  SOLSWR(M) CLOUD(M)",
        "#                    /TEQ   /RELH                      /HTCOEF",
    };
    for (const char *l : hdr) fprintf(fp, "%s\n", l);
    const double cvt = 1 / (data_basetime * 1000.0);   // mm per interval -> m/s
    fprintf(fp, "      %d %lf %lf %d %E %E %lf %lf\n", ntime, efdc_basetime, 1.0, 1, cvt, cvt, 1.0, 1.0);
    const double *p = thermal;   // IASWRAD REVC RCHC SWRATNF SWRATNS FSWRATF DABEDT TBEDIT HTBED1 HTBED2
    fprintf(fp, "      %d %lf %lf %lf %lf %E %lf %lf %E %E\n", (int)p[0], p[1], p[2], p[3], p[4], p[5],
            p[6], p[7], p[8], p[9]);
    for (int t = 0; t < ntime; t++) {
        fprintf(fp, "      %lf ", ((double)t * data_basetime) / efdc_basetime);
        for (int c = 0; c < 7; c++) fprintf(fp, c == 6 ? "%lf\n" : "%lf ", rows7[7 * (size_t)t + c]);
    }
    return true;
}

bool write_show(const std::string &path, const double *v)
{
    File o(path, "w");
    if (!o) return false;
    FILE *fp = o.fp;
    fprintf(fp, "*  SHOW.INP FILE, IN FREE FORMAT ACROSS LINE\n*\n");
    fprintf(fp, "*  NSTYPE    NSHOWR    ISHOWC   JSHOWC    NSHFREQ\n*\n");
    fprintf(fp, "*  ZSSMIN    ZSSMAX    SSALMAX\n*\n");
    fprintf(fp, "    %d    %d    %d    %d    %d\n", (int)v[0], (int)v[1], (int)v[2], (int)v[3], (int)v[4]);
    fprintf(fp, "    %d    %d    %d\n", (int)v[5], (int)v[6], (int)v[7]);
    return true;
}

}  // namespace sw2dio
