// sw2d_driver.cpp -- sw2d-efdc: the drop-in executable for the reference's
// SW2D run (main() of SW2D-EFDC.cu, 3682-5166) on the MI355X library.
//
// For every case of db/dir.txt it reads the case files (3733-4460), runs the
// reference's time loop (4789-4877) through the C ABI of libsw2d, and writes
// the reference's outputs at the reference's output steps (4880-5072):
// SW2D/output/Results_<k>.vtk, SW2D/output/WaterLevel.txt,
// SW2D/output/TimeSimu_<k>.txt and, for coupled cases, the lake-inflow series
// EFDC/input/QSER.INP (couple_BC_efdc + write_efdc_QSER, with the
// reference's one-output lag of the velocities).  With the coupling on it
// also writes EFDC's series and initial-condition inputs (WSER, ASER, SHOW,
// DYE/SALT/TEMP.INP, DSER/TSER/SSER), the grid generator's inputs
// (EFDC/input/gefdc/: CELL.INP, gridext.inp, depdat.inp, gefdc.inp), EFDC's
// domain decomposition EFDC/input/LORP.INP and EFDC/input/EFDC.INP: a copy of EFDC/input/template/EFDC.INP made before
// the run (4281-4289) whose cards C6-C11, C14, C22-C26, C71-C73, C84 and C87
// are rewritten with the first output (4924-4933).  Running the grid
// generator and the mpirun of the Fortran lake model are not part of this
// build.
//
// usage: sw2d-efdc [--root DIR] [--kernel fused|plain] [--precision fp64|fp32|mixed]
//                  [--device N] [--autotune] [--parse-only] [--efdc-inputs]
//        sw2d-efdc --lorp CELL.INP NPROC TYPE OUT
// --efdc-inputs writes the EFDC inputs known before the run (WSER, ASER,
// SHOW, DYE/SALT/TEMP.INP, EFDC.INP, the grid generator's inputs, LORP.INP)
// and stops; no GPU is used.  --lorp writes only the decomposition of a
// CELL.INP for NPROC ranks and LORP type TYPE (1 h, 2 v, 3 r) to OUT.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <atomic>
#include <condition_variable>
#include <ctime>
#include <deque>
#include <functional>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sw2d.h"
#include "sw2d_io.hpp"

namespace {

// Output writer thread.  Formatting the text outputs (VTK with "%f" per value,
// WaterLevel.txt, QSER.INP) takes longer than the GPU needs for an output
// interval on large grids, so the main thread hands each output to this
// thread, with copies of the fields it writes, and goes on queueing the next
// iterations.  Jobs run one at a time in submission order (the files come out
// exactly as if written inline); at most `depth` jobs wait.
class AsyncWriter {
public:
    explicit AsyncWriter(size_t depth) : depth_(depth), thr_([this] { run(); }) {}
    ~AsyncWriter() { finish(); }
    void submit(std::function<void()> job)
    {
        std::unique_lock<std::mutex> lk(mu_);
        space_.wait(lk, [&] { return q_.size() < depth_; });
        q_.push_back(std::move(job));
        work_.notify_one();
    }
    // run every queued job, then stop the thread
    void finish()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            work_.notify_all();
        }
        if (thr_.joinable()) thr_.join();
    }

private:
    void run()
    {
        for (;;) {
            std::function<void()> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                work_.wait(lk, [&] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                job = std::move(q_.front());
                q_.pop_front();
                space_.notify_one();
            }
            job();
        }
    }
    size_t depth_;
    std::mutex mu_;
    std::condition_variable work_, space_;
    std::deque<std::function<void()>> q_;
    bool stop_ = false;
    std::thread thr_;  // last: starts after the members above exist
};

struct Options {
    std::string root = ".";
    int kernel = SW2D_KERNEL_FUSED;
    int precision = SW2D_FP64;
    int device = 0;
    bool autotune = false;
    bool parse_only = false;
    bool efdc_inputs = false;
    // --lorp CELL.INP NPROC TYPE OUT: only the domain decomposition file
    std::vector<std::string> lorp;
    // fault injection for tests: once this many outputs are out, the next
    // step request asks for a negative iteration count, which the library
    // rejects as it would any failed step (-1: off)
    long fail_after_outputs = -1;
};

bool parse_args(int argc, char **argv, Options &o)
{
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&](std::string &v) {
            if (i + 1 >= argc) return false;
            v = argv[++i];
            return true;
        };
        std::string v;
        if (a == "--root" && next(v)) o.root = v;
        else if (a == "--kernel" && next(v)) o.kernel = (v == "plain") ? SW2D_KERNEL_PLAIN : SW2D_KERNEL_FUSED;
        else if (a == "--precision" && next(v))
            o.precision = (v == "fp32") ? SW2D_FP32 : (v == "mixed") ? SW2D_MIXED : SW2D_FP64;
        else if (a == "--device" && next(v)) o.device = atoi(v.c_str());
        else if (a == "--autotune") o.autotune = true;
        else if (a == "--parse-only") o.parse_only = true;
        else if (a == "--efdc-inputs") o.efdc_inputs = true;
        else if (a == "--lorp" && i + 4 < argc) {
            o.lorp.assign(argv + i + 1, argv + i + 5);
            i += 4;
        }
        else if (a == "--fail-step-after-outputs" && next(v)) o.fail_after_outputs = atol(v.c_str());
        else {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return false;
        }
    }
    return true;
}

// std::ostream formatting of round(x), as the reference builds file names
std::string tempo_name(double x)
{
    std::stringstream out;
    out << x;
    return out.str();
}

void print_parse_summary(const sw2dio::Case &c, double time0, double timmax)
{
    const auto &s = c.setup;
    printf("{\"dir\": \"%s\", \"rows\": %d, \"cols\": %d, \"cellsize\": %.17g, \"nodata\": %.17g, "
           "\"xll\": %.17g, \"yll\": %.17g, \"dt\": %.17g, \"dpout\": %.17g, \"dkout\": %.17g, "
           "\"dtoq\": %.17g, \"manning\": %.17g, \"INT\": %.17g, \"INF\": %.17g, \"LWL\": %.17g, "
           "\"EV_WL_min\": %.17g, \"evaporation\": %d, \"outputs\": [%d, %d, %d, %d, %d], "
           "\"efdc\": %d, \"efdc_basetime\": %.17g, \"nlayers\": %d, \"time0\": %.17g, "
           "\"timmax\": %.17g, \"nst\": %zu, \"nq\": %d, \"dtrain\": %.17g, \"nrain\": %zu, "
           "\"rain_sum\": %.17g, \"n_out\": %zu, \"z_sum\": %.17g, \"h0_sum\": %.17g, "
           "\"mask_lake\": %ld}\n",
           c.dir.c_str(), c.dem.rows, c.dem.cols, c.dem.cellsize, c.dem.nodata, c.dem.xll, c.dem.yll,
           s.dt, s.dpout, s.dkout, s.dtoq, s.manning, s.INT, s.INF, s.LWL, s.EV_WL_min,
           s.evaporation_on, s.out_vx, s.out_vy, s.out_elev, s.out_depth, s.out_outlet, s.efdc,
           s.efdc_basetime, s.nlayers, time0, timmax, c.src_x.size(), c.nq, c.dtrain, c.rain.size(),
           [&] { double a = 0; for (double r : c.rain) a += r; return a; }(), c.out_x.size(),
           [&] { double a = 0; for (double z : c.dem.v) if (z != c.dem.nodata) a += z; return a; }(),
           [&] { double a = 0; for (double h : c.h0) a += h; return a; }(),
           [&] { long n = 0; for (int m : c.mask) n += (m == 1); return n; }());
}

// EFDC inputs the coupled run writes before stepping (4405-4425 wind,
// 4588-4612 initial constituent fields, 4767-4785 atmospheric series).
// Series shorter than the run's EFDC record count end the run with a
// message (the reference reads past them).
// byte copy of the card template (4281-4289)
bool copy_file(const std::string &from, const std::string &to)
{
    FILE *in = fopen(from.c_str(), "rb");
    if (!in) {
        printf("unknown file - %s\n", from.c_str());
        return false;
    }
    FILE *out = fopen(to.c_str(), "wb");
    bool ok = out != nullptr;
    char buf[65536];
    size_t n;
    while (ok && (n = fread(buf, 1, sizeof buf, in)) > 0) ok = fwrite(buf, 1, n, out) == n;
    ok = ok && !ferror(in);
    fclose(in);
    if (out && fclose(out) != 0) ok = false;
    if (!ok) printf(" cannot write %s\n", to.c_str());
    return ok;
}

// The EFDC.INP card values known before the run (4288-4334, 4296-4307 and
// setup.dat's EFDC block); the coupling records are added with the first
// output.  water = lake cells of mask.dat.
sw2dio::EfdcConfig efdc_config(const sw2dio::Case &c, double timmax, int water)
{
    const sw2dio::Setup &s = c.setup;
    const std::vector<double> &b = s.efdc_block;
    auto I = [&](int k) { return (int)b[k]; };
    sw2dio::EfdcConfig e;
    e.nlayers = s.nlayers;
    e.ncols = c.dem.cols;
    e.nrows = c.dem.rows;
    e.nwcells = water;
    e.TCON = s.efdc_basetime;
    e.TREF = s.efdc_basetime;
    e.TBEGIN = s.TBEGIN;
    e.NTC = (int)(timmax / s.efdc_basetime);
    e.NTSPTC = s.NTSPTC;
    e.NTCPP = e.NTC;
    e.NTCVB = e.NTC;
    e.ISPPH = I(sw2dio::kISPPH); e.NPPPH = I(sw2dio::kNPPPH);
    e.ISRPPH = I(sw2dio::kISRPPH); e.IPPHXY = I(sw2dio::kIPPHXY);
    e.ISVPH = I(sw2dio::kISVPH); e.NPVPH = I(sw2dio::kNPVPH);
    e.ISRVPH = I(sw2dio::kISRVPH); e.IVPHXY = I(sw2dio::kIVPHXY);
    e.ISTMSR = s.out_outlet;
    e.MLTMSR = (int)c.efdc_out_x.size();
    e.NBTMSR = 1;
    e.NSTMSR = e.NTC * e.NTSPTC;
    e.NWTMSR = e.NTSPTC;
    e.NTSSTSP = 1;
    e.TCTMSR = c.efdc_tctmsr;
    // (I, J) of a location: cells from the DEM's lower-left corner (4331-4332)
    const double res = c.dem.cellsize;
    for (size_t i = 0; i < c.efdc_out_x.size(); i++)
        e.LTS.push_back({(int)round(fabs((c.dem.xll - c.efdc_out_x[i]) / res)),
                         (int)round(fabs((c.dem.yll - c.efdc_out_y[i]) / res))});
    e.NTSSSS = I(sw2dio::kNTSSSS); e.MTSP = I(sw2dio::kMTSP); e.MTSC = I(sw2dio::kMTSC);
    e.MTSA = I(sw2dio::kMTSA); e.MTSUE = I(sw2dio::kMTSUE); e.MTSUT = I(sw2dio::kMTSUT);
    e.MTSU = I(sw2dio::kMTSU);
    e.ZBRADJ = b[sw2dio::kZBRADJ]; e.ZBRCVRT = b[sw2dio::kZBRCVRT]; e.HMIN = b[sw2dio::kHMIN];
    e.HADJ = b[sw2dio::kHADJ]; e.HCVRT = b[sw2dio::kHCVRT]; e.HDRY = b[sw2dio::kHDRY];
    e.HWET = b[sw2dio::kHWET]; e.BELADJ = b[sw2dio::kBELADJ]; e.BELCVRT = b[sw2dio::kBELCVRT];
    e.NWSER = I(sw2dio::kNWSER);
    e.NASER = I(sw2dio::kNASER);
    e.DX = res;
    e.DY = res;
    e.dye_on = I(sw2dio::kDyeOn);
    e.salt_on = I(sw2dio::kSaltOn);
    e.temp_on = I(sw2dio::kTempOn);
    e.temp_opt = 1;
    return e;
}

// The grid generator's inputs (make_gefdc_inputs at 4472: CELL.INP,
// gridext.inp, depdat.inp, gefdc.inp) from mask.dat, the smoothed bed and
// the case's grid_make_input.dat, into EFDC/input/gefdc/, the folder to run
// the grid generator in (the reference writes them to its working directory
// and runs it there, 4472-4535).
bool write_gefdc(const sw2dio::Case &c, const std::vector<double> &z)
{
    sw2dio::GefdcParams gp;
    std::string err;
    if (!sw2dio::read_grid_make_input(c.dir + "/grid_make_input.dat", gp, &err)) {
        printf("%s\n", err.c_str());
        return false;
    }
    const std::string dir = c.dir + "/EFDC/input/gefdc/";
    mkdir(dir.c_str(), 0755);
    int nd = 0;
    if (!sw2dio::write_gefdc_inputs(dir, c.mask_hdr.cols, c.mask_hdr.rows, c.mask_hdr.cellsize, c.mask,
                                    z, gp, &nd, &err)) {
        printf(" %s\n", err.c_str());
        return false;
    }
    // the domain decomposition for setup.dat's number of EFDC ranks and LORP
    // type (the reference runs its partitioner on CELL.INP and renames the
    // chosen file, 4480-4507)
    const std::vector<double> &b = c.setup.efdc_block;
    const std::vector<unsigned char> types =
        sw2dio::gefdc_cell_types(c.mask_hdr.cols, c.mask_hdr.rows, c.mask);
    if (!sw2dio::write_lorp(c.dir + "/EFDC/input/LORP.INP", types, c.mask_hdr.cols, c.mask_hdr.rows,
                            (int)b[sw2dio::kNProcessors], (int)b[sw2dio::kLorpType], &err)) {
        printf(" %s\n", err.c_str());
        return false;
    }
    return true;
}

// records come back from sw2d_couple_bc as S, N, E, W; QSER.INP and C24
// list them N, S, E, W
const int kind_of_edge[4] = {1, 0, 2, 3};

// EFDC/input/EFDC.INP's cards with the coupling records of the first output
// (4924-4933); I/J/counts as sw2d_couple_bc returns them, lists cap apart
bool write_efdc_cards(const sw2dio::Case &c, sw2dio::EfdcConfig cfg, const int *I, const int *J,
                      const int *counts, int cap)
{
    int n = 0;
    for (int e = 0; e < 4; e++) {
        const int k = kind_of_edge[e];
        cfg.bc[e].clear();
        for (int i = 0; i < counts[k]; i++)
            cfg.bc[e].push_back({I[(size_t)k * cap + i], J[(size_t)k * cap + i]});
        n += counts[k];
    }
    cfg.nqsij = cfg.nqser = n;
    std::string err;
    if (!sw2dio::write_efdc_config(c.dir + "/EFDC/input/EFDC.INP", cfg, &err)) {
        printf(" EFDC.INP: %s\n", err.c_str());
        return false;
    }
    return true;
}

bool write_efdc_setup_files(const sw2dio::Case &c, const std::vector<double> &ev, double time0,
                            double timmax, long lkout)
{
    const sw2dio::Setup &s = c.setup;
    const std::vector<double> &blk = s.efdc_block;
    const std::string in = c.dir + "/EFDC/input/";
    const int R = c.dem.rows, C = c.dem.cols, nl = s.nlayers;
    const int ntime = (int)(((timmax - time0) / s.dt) / (double)lkout);   // efdc_ntimes (4241)
    auto wrote = [&](bool done, const char *file) {
        if (!done) printf(" cannot write %s%s\n", in.c_str(), file);
        return done;
    };
    auto short_series = [&](const char *what, size_t have, size_t need) {
        printf(" %s holds %zu values, %zu needed for %d EFDC records\n", what, have, need, ntime);
        return false;
    };
    if ((int)blk[sw2dio::kNWSER] == 1) {
        if (c.wind.size() < 2 * (size_t)std::max(ntime, 0))
            return short_series("wind_velocities.dat", c.wind.size() / 2, (size_t)ntime);
        if (!wrote(sw2dio::write_wser(in + "WSER.INP", c.wind, c.dtrain, s.efdc_basetime, ntime), "WSER.INP"))
            return false;
    }
    int water = 0;
    for (int m : c.mask) water += (m == 1);
    if ((int)blk[sw2dio::kDyeOn] == 1 &&
        !wrote(sw2dio::write_initial_field(in + "DYE.INP", C, R, nl, water, c.mask,
                                           [&](size_t, int) { return blk[sw2dio::kDyeInit]; }), "DYE.INP"))
        return false;
    if ((int)blk[sw2dio::kSaltOn] == 1 &&
        !wrote(sw2dio::write_initial_field(in + "SALT.INP", C, R, nl, water, c.mask,
                                           [&](size_t, int) { return blk[sw2dio::kSaltInit]; }), "SALT.INP"))
        return false;
    if ((int)blk[sw2dio::kTempOn] == 1) {
        // temperature decreasing with depth below the surface value: layer l
        // of the initial depth column sits (nl - 1 - l) / nl of it deep
        const double tb = blk[sw2dio::kTempBot], ts = blk[sw2dio::kTempSurf];
        auto temp = [&](size_t m, int l) {
            const double L_DEP = (c.h0[m] / nl) * ((nl - 1) - l);
            return tb + (ts - tb) / (1.00 + L_DEP);
        };
        if (!wrote(sw2dio::write_initial_field(in + "TEMP.INP", C, R, nl, water, c.mask, temp), "TEMP.INP"))
            return false;
    }
    if (s.evaporation_on == 1) {
        const size_t n = (size_t)std::max(ntime, 0);
        if (c.patm.size() < n) return short_series("atm_press.dat", c.patm.size(), n);
        if (c.relh.size() < n) return short_series("rel_humidity.dat", c.relh.size(), n);
        if (c.T.size() < n || c.rain.size() < n || c.Rg.size() < n || ev.size() < n)
            return short_series("temperatures/rain/solar_radiation.dat",
                                std::min({c.T.size(), c.rain.size(), c.Rg.size(), ev.size()}), n);
        std::vector<double> rows(7 * n);
        for (size_t t = 0; t < n; t++) {
            double *r = &rows[7 * t];
            r[0] = c.patm[t];
            r[1] = c.T[t];
            r[2] = c.relh[t];
            r[3] = (1 - s.LWL) * c.rain[t];
            r[4] = ev[t];
            r[5] = c.Rg[t];
            r[6] = 0.00;   // cloud cover
        }
        if (!wrote(sw2dio::write_aser(in + "ASER.INP", rows, &blk[sw2dio::kIASWRAD], c.dtrain,
                                      s.efdc_basetime, ntime), "ASER.INP"))
            return false;
    }
    return true;
}

int run_case(const Options &opt, const std::string &name)
{
    printf("%s\n", " ******************************************************************** ");
    printf("CASE: %s\n", name.c_str());
    sw2dio::Case c;
    const std::string err = sw2dio::read_case(opt.root, name, c);
    if (!err.empty()) {
        printf("%s\n", err.c_str());
        return 1;
    }
    const sw2dio::Setup &s = c.setup;
    const int R = c.dem.rows, C = c.dem.cols;
    const double dt = s.dt;
    // output cadence and time window (4187-4188, 4230-4235)
    long lpout = lround(s.dpout / dt), lkout = lround(s.dkout / dt);
    if (lpout <= 0) lpout = 1;
    if (lkout <= 0) lkout = 1;
    const double time0 = 3600.0 * 24.0 * s.tday0 + 3600.0 * s.thour0 + 60.0 * s.tmin0 + s.tsec0;
    double timmax = 3600.0 * 24.0 * s.tday + 3600.0 * s.thour + 60.0 * s.tmin + s.tsec;
    if ((s.tday == 0) & (s.thour == 0) & (s.tmin == 0) & (s.tsec == 0))
        timmax = ((double)c.rain.size() - 1) * c.dtrain;  // (cont_rain - 2) * dtrain
    if (opt.parse_only) {
        print_parse_summary(c, time0, timmax);
        return 0;
    }
    if (opt.efdc_inputs) {
        if (s.efdc != 1) {
            printf(" the EFDC coupling is off in setup.dat: no EFDC inputs to write\n");
            return 1;
        }
        std::vector<double> ev;
        if (s.evaporation_on == 1) {
            ev.resize(c.T.size());
            sw2d_evaporation_series(c.albedo, c.T.data(), c.Rg.data(), (int)c.T.size(), c.dtrain, ev.data());
        }
        if (!write_efdc_setup_files(c, ev, time0, timmax, lkout)) return 1;
        if (!sw2dio::write_show(c.dir + "/EFDC/input/SHOW.INP", &s.efdc_block[sw2dio::kNSTYPE])) {
            printf(" cannot write %s/EFDC/input/SHOW.INP\n", c.dir.c_str());
            return 1;
        }
        std::vector<double> zs = c.dem.v;
        sw2d_lake_catchment_smooth(R, C, c.mask.data(), zs.data());  // 4460
        if (!write_gefdc(c, zs)) return 1;
        // EFDC.INP: the first output's coupling records depend only on
        // mask.dat (couple_BC_efdc at output 0 takes the initial depths and
        // zero velocities, 4902), so its cards are known before the run
        if (!copy_file(c.dir + "/EFDC/input/template/EFDC.INP", c.dir + "/EFDC/input/EFDC.INP"))
            return 1;
        const size_t RC = (size_t)R * C;
        const int cap = (int)RC;
        std::vector<double> zero(RC, 0.0), q(4 * RC);
        std::vector<int> cells(4 * RC), I(4 * RC), J(4 * RC), counts(4);
        if (sw2d_couple_bc(R, C, c.dem.cellsize, c.dem.cellsize, c.mask.data(), c.h0.data(), zero.data(),
                           zero.data(), cap, cells.data(), I.data(), J.data(), q.data(),
                           counts.data()) != SW2D_OK) {
            printf(" sw2d_couple_bc failed: %s\n", sw2d_last_error(nullptr));
            return 1;
        }
        int water = 0;
        for (int m : c.mask) water += (m == 1);
        if (!write_efdc_cards(c, efdc_config(c, timmax, water), I.data(), J.data(), counts.data(), cap))
            return 1;
        return 0;
    }

    std::vector<double> z = c.dem.v;
    if (s.efdc == 1) {
        sw2d_lake_catchment_smooth(R, C, c.mask.data(), z.data());  // 4460
        if (!write_gefdc(c, z)) return 1;
    }

    sw2d_grid_desc grid{R, C, c.dem.xll, c.dem.yll, c.dem.cellsize, c.dem.nodata};
    sw2d_params prm;
    sw2d_default_params(&prm);
    prm.g = 9.80;
    prm.manning = s.manning;
    prm.th = 1.0e-4;
    prm.dt = dt;
    prm.time0 = time0;
    prm.INT = s.INT; prm.INF = s.INF; prm.LWL = s.LWL; prm.EV_WL_min = s.EV_WL_min;
    prm.kernel = opt.kernel;
    prm.precision = opt.precision;
    prm.device = opt.device;
    sw2d_handle *h = nullptr;
    if (sw2d_create(&grid, z.data(), c.h0.data(), &prm, &h) != SW2D_OK) {
        printf("sw2d_create failed: %s\n", sw2d_last_error(nullptr));
        return 1;
    }
    auto check = [&](int rc, const char *what) {
        if (rc != SW2D_OK) {
            printf("%s failed: %s\n", what, sw2d_last_error(h));
            return false;
        }
        return true;
    };
    bool ok = check(sw2d_set_rain_series(h, c.rain.data(), (int)c.rain.size(), c.dtrain), "rain");
    ok = ok && check(sw2d_set_sources(h, c.src_x.data(), c.src_y.data(), (int)c.src_x.size(),
                                      c.Q.data(), c.nq, s.dtoq), "sources");
    std::vector<double> ev;
    if (ok && s.evaporation_on == 1) {
        ev.resize(c.T.size());
        sw2d_evaporation_series(c.albedo, c.T.data(), c.Rg.data(), (int)c.T.size(), c.dtrain, ev.data());
        ok = check(sw2d_set_evaporation(h, ev.data(), (int)ev.size()), "evaporation");
    }
    if (ok && s.efdc == 1) ok = write_efdc_setup_files(c, ev, time0, timmax, lkout);
    if (ok && s.efdc == 1)
        ok = copy_file(c.dir + "/EFDC/input/template/EFDC.INP", c.dir + "/EFDC/input/EFDC.INP");
    if (ok && opt.autotune) {
        int H = 0;
        ok = check(sw2d_autotune(h, 10, &H), "autotune");
    }
    if (!ok) {
        sw2d_destroy(h);
        return 1;
    }

    const std::string outdir = c.dir + "/SW2D/output/";
    FILE *wl = fopen((outdir + "WaterLevel.txt").c_str(), "w");
    if (!wl) {
        printf("cannot write %sWaterLevel.txt\n", outdir.c_str());
        sw2d_destroy(h);
        return 1;
    }
    if (s.out_outlet != 1) fprintf(wl, "%s\n", "Outlet off");
    const double res = c.dem.cellsize;
    std::vector<double> outx(c.out_x.size()), outy(c.out_y.size());
    for (size_t i = 0; i < outx.size(); i++) {
        outx[i] = c.out_x[i] / res;
        outy[i] = c.out_y[i] / res;
    }

    // coupling state (4241-4403): records, their values per output, lagged velocities
    const size_t RC = (size_t)R * C;
    sw2dio::QserSeries qser;
    std::vector<double> uua_host(RC, 0.0), vva_host(RC, 0.0);
    const int cap = (int)RC;
    std::vector<int> cells, I, J, counts(4);
    std::vector<double> q;
    if (s.efdc == 1) {
        cells.resize(4 * RC); I.resize(4 * RC); J.resize(4 * RC); q.resize(4 * RC);
    }
    sw2dio::EfdcConfig efdc_cards;
    if (s.efdc == 1) {
        int water = 0;
        for (int m : c.mask) water += (m == 1);
        efdc_cards = efdc_config(c, timmax, water);
    }

    AsyncWriter writer(2);
    std::atomic<bool> write_failed{false};  // set by the writer thread
    // fields an output needs; each output is a snapshot (device copy, then a
    // pinned-memory copy that overlaps the following steps) that the writer
    // thread reads and releases
    unsigned fields = 0;
    if (s.efdc == 1 || s.out_depth == 1) fields |= 1u << SW2D_FIELD_H;
    if (s.efdc == 1) fields |= (1u << SW2D_FIELD_UUA) | (1u << SW2D_FIELD_VVA);
    if (s.out_vx == 1) fields |= 1u << SW2D_FIELD_VV;
    if (s.out_vy == 1) fields |= 1u << SW2D_FIELD_UU;
    double time = time0;
    long mstep = 0;
    int tq = -1;
    long out0 = -1;
    long pending = 0;
    std::string tempo = "0";
    bool run_failed = false;   // a step or snapshot failed: the outputs stop there
    auto step_request = [&](long n) {
        const bool inject = opt.fail_after_outputs >= 0 && tq + 1 >= opt.fail_after_outputs;
        return check(sw2d_step(h, inject ? -1 : (int)n), "step");
    };
    const std::clock_t start = std::clock();
    while (time + dt <= timmax) {
        out0 = out0 + 1;
        if (mstep % lpout == 0) printf(" %d\n", (int)time);
        pending++;
        time = time + dt;
        mstep = mstep + 1;
        time = time + dt;
        mstep = mstep + 1;
        if (!((mstep % lkout == 0) || (out0 == 0))) continue;
        if (!step_request(pending)) {
            run_failed = true;
            break;
        }
        pending = 0;
        tq = tq + 1;
        tempo = (out0 > 0) ? tempo_name(round(mstep * dt / s.dkout)) : std::string("0");
        long ticket = -1;
        if (fields && !check(sw2d_snapshot(h, fields, 3600.0, &ticket), "snapshot")) {
            run_failed = true;
            break;
        }
        if (s.efdc == 1 && tq == 0) {
            // written with the first output, beside the EFDC.INP card update (4934-4936)
            const double *sv = &s.efdc_block[sw2dio::kNSTYPE];
            if (!sw2dio::write_show(c.dir + "/EFDC/input/SHOW.INP", sv)) write_failed = true;
        }
        // the coupling state (records, QSER series, lagged velocities) and the
        // files belong to the writer thread; jobs run in output order
        writer.submit([&, ticket, tq, out0, efdc_time = mstep * dt / s.efdc_basetime,
                       path = outdir + "Results_" + tempo + ".vtk"] {
            auto field = [&](int f) -> const double * {
                const double *p = nullptr;
                if (sw2d_snapshot_field(h, ticket, f, &p) != SW2D_OK) write_failed = true;
                return p;
            };
            const double *hnow = (fields >> SW2D_FIELD_H & 1u) ? field(SW2D_FIELD_H) : nullptr;
            if (s.efdc == 1) {
                const double *hh = (tq == 0) ? c.h0.data() : hnow;
                const double tt = (tq == 0) ? 0.0000 : efdc_time;
                const double *ua = field(SW2D_FIELD_UUA), *va = field(SW2D_FIELD_VVA);
                if (hh && ua && va) {
                    sw2d_couple_bc(R, C, res, res, c.mask.data(), hh, vva_host.data(), uua_host.data(), cap,
                                   cells.data(), I.data(), J.data(), q.data(), counts.data());
                    for (int e = 0; e < 4; e++) {
                        const int k = kind_of_edge[e];
                        qser.nrec[e] = counts[k];
                        qser.time[e].push_back(counts[k] > 0 ? tt : 0.0);
                        qser.val[e].emplace_back(q.begin() + (size_t)k * cap,
                                                 q.begin() + (size_t)k * cap + counts[k]);
                    }
                    if (tq == 0 && !write_efdc_cards(c, efdc_cards, I.data(), J.data(), counts.data(), cap))
                        write_failed = true;
                    // the velocities used by the next output are this
                    // output's, taken after the coupling call (4913-4916)
                    std::copy(ua, ua + RC, uua_host.begin());
                    std::copy(va, va + RC, vva_host.begin());
                    if (!sw2dio::write_qser(c.dir + "/EFDC/input/QSER.INP", 0, s.nlayers, qser,
                                            s.efdc_basetime, tq))
                        write_failed = true;
                }
            }
            const double *depth = (out0 == 0) ? c.h0.data() : hnow;
            if (s.out_depth == 1 && s.out_outlet == 1 && depth)
                sw2dio::write_water_level(wl, R, C, c.dem.xll / res, c.dem.yll / res, outx, outy, depth);
            sw2dio::VtkFields vf;
            vf.depth = (s.out_depth == 1) ? depth : nullptr;
            vf.vv = (s.out_vx == 1) ? field(SW2D_FIELD_VV) : nullptr;
            vf.uu = (s.out_vy == 1) ? field(SW2D_FIELD_UU) : nullptr;
            vf.z = (s.out_elev == 1) ? z.data() : nullptr;
            if (!sw2dio::write_vtk(path, C, R, res, res, c.dem.xll, c.dem.yll, vf)) write_failed = true;
            if (ticket >= 0) sw2d_snapshot_release(h, ticket);
        });
    }
    if (!run_failed && pending > 0 && !step_request(pending)) run_failed = true;
    if (sw2d_sync(h) != SW2D_OK) run_failed = true;
    writer.finish();
    if (run_failed) {
        printf(" the run stopped at t = %.6f s after a library error: %d output(s) written, "
               "the rest of the run is missing\n", time, tq + 1);
        fclose(wl);
        sw2d_destroy(h);
        return 1;
    }
    // constituent series of the coupling records, after the run (5076-5092)
    if (s.efdc == 1 && tq >= 0) {
        const std::string in = c.dir + "/EFDC/input/";
        const int nl = s.nlayers;
        const double base = s.efdc_basetime;
        const std::vector<double> &blk = s.efdc_block;
        if ((int)blk[sw2dio::kDyeOn] == 1 &&
            !sw2dio::write_constituent_series(in + "DSER.INP", 'D', 0, nl, qser,
                                              [&](int) { return blk[sw2dio::kDyeIn]; }, base, tq))
            write_failed = true;
        if ((int)blk[sw2dio::kTempOn] == 1) {
            // one temperature per output from temperatures.dat
            if ((int)c.T.size() <= tq) {
                printf(" TSER.INP needs %d values of temperatures.dat (evaporation on), %zu read\n",
                       tq + 1, c.T.size());
                write_failed = true;
            } else if (!sw2dio::write_constituent_series(in + "TSER.INP", 'T', 0, nl, qser,
                                                         [&](int t) { return c.T[t]; }, base, tq)) {
                write_failed = true;
            }
        }
        if ((int)blk[sw2dio::kSaltOn] == 1 &&
            !sw2dio::write_constituent_series(in + "SSER.INP", 'S', 0, nl, qser,
                                              [&](int) { return blk[sw2dio::kSaltIn]; }, base, tq))
            write_failed = true;
    }
    if (write_failed) {
        printf(" an output file could not be written (VTK, QSER.INP, SHOW.INP or a constituent series)\n");
        sw2d_destroy(h);
        return 1;
    }
    const double duration = (std::clock() - start) / (double)CLOCKS_PER_SEC;
    printf("printf: %g\n", duration);
    FILE *ts = fopen((outdir + "TimeSimu_" + tempo + ".txt").c_str(), "w");
    if (ts) {
        fprintf(ts, " %lf\n", duration);
        fclose(ts);
    }
    fclose(wl);
    if (s.efdc == 1) printf(" the EFDC lake model is not run by this build (its inputs are in EFDC/input)\n");
    sw2d_destroy(h);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    Options opt;
    if (!parse_args(argc, argv, opt)) return 2;
    if (!opt.lorp.empty()) {
        std::vector<unsigned char> types;
        int cols = 0, rows = 0;
        std::string err;
        if (!sw2dio::read_cell_types(opt.lorp[0], types, cols, rows, &err) ||
            !sw2dio::write_lorp(opt.lorp[3], types, cols, rows, atoi(opt.lorp[1].c_str()),
                                atoi(opt.lorp[2].c_str()), &err)) {
            printf("%s\n", err.c_str());
            return 1;
        }
        return 0;
    }
    std::vector<std::string> names;
    const std::string err = sw2dio::read_dir_list(opt.root, names);
    if (!err.empty()) {
        printf("%s\n", err.c_str());
        return 1;
    }
    int rc = 0;
    for (const std::string &n : names) rc |= run_case(opt, n);
    return rc;
}
