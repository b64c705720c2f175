// sw2d_efdc_config.cpp -- the EFDC.INP card update the reference makes with
// its first output (write_efdc_CONFIG, SW2D-EFDC.cu:1069-2024, called at
// 4924-4937 on the copy of EFDC/input/template/EFDC.INP made at 4281-4289).
//
// Seventeen cards are replaced, in this order: C6 C7 C8 C9 C10 C11 C14 C22
// C23 C24 C25 C26 C71 C72 C73 C84 C87.  For card Cn with successor card Cm
// the reference finds
//   - the start: the first whitespace-delimited token "Cn" of the file, and
//     the position just past the token before it and the one character
//     after that (fscanf "%s%*c"); with CRLF lines that is the LF before
//     "Cn", otherwise "Cn" itself;
//   - the kept tail: what follows the first token "Cm", its next character
//     and the whitespace after that (fscanf "%s%*c\n"),
// and writes, from the start: a newline (some cards), the card's comment
// block, header and data lines, a dashed line, "Cm " and the kept tail.
// The card texts and printf layouts below are the file format EFDC reads
// (its SEEK finds a card's title line and then its header line, and reads
// the data after it), so they are kept as the reference prints them.
//
// Deliberate difference: the reference rewrites the file in place, so when a
// card gets shorter the old file's last four bytes stay behind a run of
// spaces, and its copy of the tail drops the file's last byte (successors of
// two characters) or appends one byte it never read (four characters, C11A
// and C71A).  Here each card is spliced into the whole text and the file is
// written once: the tail is kept exactly.
#include <ctype.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "sw2d_io.hpp"

namespace sw2dio {

namespace {

const char kDashes[] = "-------------------------------------------------------------------------------";

struct Token {
    size_t begin = 0, end = 0;  // [begin, end) of the token in the text
    size_t prev_end = 0;        // where the reference's card text starts
    bool found = false;
};

// first whitespace-delimited token equal to `tok` (fscanf's %s tokenising)
Token find_token(const std::string &s, const char *tok)
{
    Token t;
    const size_t n = s.size(), len = strlen(tok);
    size_t i = 0, prev = 0;
    bool have_prev = false;
    while (i < n) {
        while (i < n && isspace((unsigned char)s[i])) i++;
        if (i >= n) break;
        const size_t b = i;
        while (i < n && !isspace((unsigned char)s[i])) i++;
        if (i - b == len && s.compare(b, len, tok) == 0) {
            t.begin = b;
            t.end = i;
            t.found = true;
            // past the previous token and the one character (%*c) after it
            t.prev_end = have_prev ? prev + 1 : 0;
            return t;
        }
        prev = i;
        have_prev = true;
    }
    return t;
}

std::string fmt(const char *f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char *f, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

std::string lines(std::initializer_list<const char *> ls)
{
    std::string out;
    for (const char *l : ls) {
        out += l;
        out += '\n';
    }
    return out;
}

// one card: its name, its successor, whether the reference prints a newline
// before it, and its text (comments, header, data)
struct Card {
    const char *name, *next;
    bool lead;
    std::string text;
};

}  // namespace

bool write_efdc_config(const std::string &path, const EfdcConfig &c, std::string *err)
{
    std::string s;
    {
        FILE *fp = fopen(path.c_str(), "rb");
        if (!fp) {
            if (err) *err = "cannot read " + path;
            return false;
        }
        char buf[65536];
        size_t got;
        while ((got = fread(buf, 1, sizeof buf, fp)) > 0) s.append(buf, got);
        fclose(fp);
    }
    const int salt = c.salt_on, temp = c.temp_on, dye = c.dye_on;
    std::vector<Card> cards;

    {   // C6 (1101-1154)
        std::string t = lines({
            "C6 DISSOLVED AND SUSPENDED CONSTITUENT TRANSPORT SWITCHES",
            "*  TURB INTENSITY=0,SAL=1,TEM=2,DYE=3,SFL=4,TOX=5,SED=6,SND=7,CWQ=8",
            "*  ",
            "*  ISTRAN:  1 OR GREATER TO ACTIVATE TRANSPORT",
            "*  ISTOPT:    NONZERO FOR TRANSPORT OPTIONS, SEE USERS MANUAL",
            "*  ISCDCA:  0 FOR STANDARD DONOR CELL UPWIND DIFFERENCE ADVECTION (3TL ONLY)",
            "*           1 FOR CENTRAL DIFFERENCE ADVECTION FOR THREE TIME LEVEL STEPS (3TL ONLY)",
            "*           2 FOR EXPERIMENTAL UPWIND DIFFERENCE ADVECTION (FOR RESEARCH) (3TL ONLY)",
            "*  ISADAC:  1 TO ACTIVATE ANTI-NUMERICAL DIFFUSION CORRECTION TO",
            "*             STANDARD DONOR CELL SCHEME",
            "*  ISFCT:   1 TO ADD FLUX LIMITING TO ANTI-NUMERICAL DIFFUSION CORRECTION",
            "*  ISPLIT:  1 TO OPERATOR SPLIT HORIZONTAL AND VERTICAL ADVECTION",
            "*             (FOR RESEARCH PURPOSES)",
            "*  ISADAH:  1 TO ACTIVATE ANTI-NUM DIFFUSION CORRECTION TO HORIZONTAL",
            "*             SPLIT ADVECTION STANDARD DONOR CELL SCHEME (FOR RESEARCH)",
            "*  ISADAV:  1 TO ACTIVATE ANTI-NUM DIFFUSION CORRECTION TO VERTICAL",
            "*             SPLIT ADVECTION STANDARD DONOR CELL SCHEME (FOR RESEARCH)",
            "*  ISCI:    1 TO READ CONCENTRATION FROM FILE restart.inp",
            "*  ISCO:    1 TO WRITE CONCENTRATION TO FILE restart.out",
            "* ",
            "C6  ISTRAN  ISTOPT  ISCDCA  ISADAC   ISFCT  ISPLIT  ISADAH  ISADAV   ISCI    ISCO"});
        // one row per constituent: ISTRAN, ISTOPT, then the advection switches
        struct Row { int istran, istopt, isadac, isfct; const char *id; };
        const Row rows[] = {{0, 0, 0, 0, "!TURB 0"}, {salt, salt, 1, 1, "!SAL 1"},
                            {temp, c.temp_opt, 1, 1, "!TEM 2"}, {dye, dye, 1, 1, "!DYE 3"},
                            {0, 0, 1, 1, "!SFL 4"}, {0, 0, 1, 1, "!TOX 5"}, {0, 0, 1, 1, "!SED 6"},
                            {0, 0, 1, 1, "!SND 7"}, {0, 0, 1, 1, "!CWQ "}};
        for (const Row &r : rows)
            t += fmt("    %d      %d      %d      %d      %d      %d      %d      %d      %d      %d     %s\n",
                     r.istran, r.istopt, 0, r.isadac, r.isfct, 0, 0, 0, 0, 0, r.id);
        cards.push_back({"C6", "C7", true, t});
    }
    {   // C7 (1196-1217)
        std::string t = lines({
            "C7 TIME-RELATED INTEGER PARAMETERS", "*",
            "*  NTC:     NUMBER OF REFERENCE TIME PERIODS IN RUN",
            "*  NTSPTC : NUMBER OF TIME STEPS PER REFERENCE TIME PERIOD",
            "*  NLTC : NUMBER OF LINEARIZED REFERENCE TIME PERIODS",
            "*  NLTC : NUMBER OF TRANSITION REF TIME PERIODS TO FULLY NONLINEAR",
            "*  NTCPP : NUMBER OF REFERENCE TIME PERIODS BETWEEN FULL PRINTED OUTPUT",
            "*           TO FILE EFDC.OUT",
            "*  NTSTBC : NUMBER OF TIME STEPS BETWEEN USING A TWO TIME LEVEL TRAPEZOIDAL",
            "*           CORRECTION TIME STEP, ** MASS BALANCE PRINT INTERVAL **",
            "*  NTCNB : NUMBER OF REFERENCE TIME PERIODS WITH NO BUOYANCY FORCING(not used)",
            "*  NTCVB : NUMBER OF REF TIME PERIODS WITH VARIABLE BUOYANCY FORCING",
            "*  NTSMMT : NUMBER OF NUMBER OF REF TIME TO AVERAGE OVER TO OBTAIN",
            "*           RESIDUAL OR MEAN MASS TRANSPORT VARIABLES",
            "*  NFLTMT : USE 1 (FOR RESEARCH PURPOSES)",
            "*  NDRYSTP : MIN NO.OF TIME STEPS A CELL REMAINS DRY AFTER INTIAL DRYING",
            "*           -NDRYSTP FOR ISDRY = -99 TO ACTIVATE WASTING WATER IN DRY CELLS",
            "C7   NTC  NTSPTC    NLTC    NTTC   NTCPP  NTSTBC   NTCNB  NTCVB  NTSMMT  NFLTMT NDRYSTP"});
        t += fmt("    %d      %d      %d      %d      %d      %d      %d      %d      %d      %d     %d\n",
                 c.NTC, c.NTSPTC, 0, 0, c.NTCPP, 0, 0, 0, c.NTCVB, 1, 16);
        cards.push_back({"C7", "C8", true, t});
    }
    {   // C8 (1244-1263)
        std::string t = lines({
            "C8 TIME-RELATED REAL PARAMETERS", "*",
            "*  TCON:     CONVERSION MULTIPLIER TO CHANGE TBEGIN TO SECONDS",
            "*  TBEGIN : TIME ORIGIN OF RUN",
            "*  TREF : REFERENCE TIME PERIOD IN sec(i.e. 44714.16S OR 86400S)",
            "*  CORIOLIS : CONSTANT CORIOLIS PARAMETER IN 1 / sec = 2 * 7.29E-5*SIN(LAT)",
            "*  ISCORV : 1 TO READ VARIABLE CORIOLIS COEFFICIENT FROM LXLY.INP FILE",
            "*  ISCCA : WRITE DIAGNOSTICS FOR MAX CORIOLIS - CURV ACCEL TO FILEEFDC.LOG",
            "*  ISCFL : 1 WRITE DIAGNOSTICS OF MAX THEORETICAL TIME STEP TO CFL.OUT",
            "*            GT 1  TIME STEP ONLY AT INTERVAL ISCFL FOR ENTIRE RUN",
            "*  ISCFLM : 1  TO MAP LOCATIONS OF MAX TIME STEPS OVER ENTIRE RUN",
            "*  DTSSFAC : DYNAMIC TIME STEPPING IF 0.0.LT.DTSSFAC.LT.1.0", "*",
            "C8  TCON  TBEGIN    TREF CORIOLIS ISCORV   ISCCA   ISCFL  ISCFLM DTSSFAC"});
        t += fmt("    %lf      %lf      %lf      %lf      %d      %d      %d      %d      %lf\n",
                 c.TCON, c.TBEGIN, c.TREF, 0.0, 0, 0, 0, 0, 0.0);
        cards.push_back({"C8", "C9", true, t});
    }
    {   // C9 (1289-1322)
        std::string t = lines({
            "C9 SPACE-RELATED AND SMOOTHING  PARAMETERS", "*",
            "*  KC:      NUMBER OF VERTICAL LAYERS",
            "*  IC : NUMBER OF CELLS IN I DIRECTION",
            "*  JC : NUMBER OF CELLS IN J DIRECTION",
            "*  LC : NUMBER OF ACTIVE CELLS IN HORIZONTAL + 2",
            "* LVC : NUMBER OF VARIABLE SIZE HORIZONTAL CELLS",
            "*  ISCO : 1 FOR CURVILINEAR - ORTHOGONAL GRID(LVC = LC - 2)",
            "*  NDM : NUMBER OF DOMAINS FOR HORIZONTAL DOMAIN DECOMPOSITION",
            "*           (NDM = 1, FOR MODEL EXECUTION ON A SINGLE PROCESSOR SYSTEM OR",
            "*             NDM = MM*NCPUS, WHERE MM IS AN INTEGER AND NCPUS IS THE NUMBER",
            "*             OF AVAILABLE CPU'S FOR MODEL EXECUTION ON A PARALLEL MULTIPLE PROCESSOR SYSTEM )",
            "*  LDM:     NUMBER OF WATER CELLS PER DOMAIN(LDM = (LC - 2) / NDM, FOR MULTIPE VECTOR PROCESSORS,",
            "*LDM MUST BE AN INTEGER MULTIPLE OF THE VECTOR LENGTH OR",
            "*             STRIDE NVEC THUS CONSTRAINING LC - 2 TO BE AN INTEGER MULTIPLE OF NVEC)",
            "*  ISMASK : 1 FOR MASKING WATER CELL TO LAND OR ADDING THIN BARRIERS",
            "*            USING INFORMATION IN FILE MASK.INP",
            "*  ISPGNS : 1 FOR IMPLEMENTING A PERIODIC GRID IN COMP N - S DIRECTION OR",
            "*            CONNECTING ARBITRATY CELLS USING INFO IN FILE MAPPGNS.INP",
            "*  NSHMAX : NUMBER OF DEPTH SMOOTHING PASSES",
            "*  NSBMAX : NUMBER OF INITIAL SALINITY FIELD SMOOTHING PASSES",
            "*  WSMH : DEPTH SMOOTHING WEIGHT",
            "*  WSMB : SALINITY SMOOTHING WEIGHT", "*", "*", "C",
            "C9   KC   IC   JC   LC  LVC ISCO  NDM  LDM  ISMASK  ISPGNS  NSHMAX  NSBMAX    WSMH    WSMB"});
        t += fmt("    %d      %d      %d      %d      %d      %d      %d      %d     %d     %d     %d     %d     %lf     %lf\n",
                 c.nlayers, c.ncols, c.nrows, c.nwcells + 2, c.nwcells, 1, 1, c.nwcells, 0, 0, 0, 0,
                 0.03125, 0.06250);
        cards.push_back({"C9", "C10", true, t});
    }
    {   // C10 (1348-1358): equal layer thicknesses
        std::string t = lines({"C10 LAYER THICKNESS IN VERTICAL", "*",
                               "*    K:  LAYER NUMBER, K = 1, KC",
                               "*  DZC : DIMENSIONLESS LAYER THICKNESS(THICKNESSES MUST SUM TO 1.0)", "*",
                               "C10  K   DZC"});
        for (int k = 0; k < c.nlayers; k++) t += fmt("    %d      %lf\n", k + 1, 1.0000 / c.nlayers);
        cards.push_back({"C10", "C11", false, t});
    }
    {   // C11 (1384-1405)
        std::string t = lines({
            "C11 GRID, ROUGHNESS AND DEPTH PARAMETERS", "*",
            "*  DX:       CARTESIAN CELL LENGTH IN X OR I DIRECTION",
            "*  DY:       CARTESION CELL LENGHT IN Y OR J DIRECTION",
            "*  DXYCVT:   MULTIPLY DX AND DY BY TO OBTAIN METERS",
            "*  IMD:      GREATER THAN 0 TO READ MODDXDY.INP FILE",
            "*  ZBRADJ:   LOG BDRY LAYER CONST OR VARIABLE ROUGH HEIGHT ADJ IN METERS",
            "*  ZBRCVRT:  LOG BDRY LAYER VARIABLE ROUGHNESS HEIGHT CONVERT TO METERS",
            "*  HMIN:     MINIMUM DEPTH OF INPUTS DEPTHS IN METERS",
            "*  HADJ:     ADJUCTMENT TO DEPTH FIELD IN METERS",
            "*  HCVRT:    CONVERTS INPUT DEPTH FIELD TO METERS",
            "*  HDRY:     DEPTH AT WHICH CELL OR FLOW FACE BECOMES DRY",
            "*  HWET:     DEPTH AT WHICH CELL OR FLOW FACE BECOMES WET",
            "*  BELADJ:   ADJUCTMENT TO BOTTOM BED ELEVATION FIELD IN METERS",
            "*  BELCVRT:  CONVERTS INPUT BOTTOM BED ELEVATION FIELD TO METERS", "*",
            "C11   DX      DY   DXYCVT    IMD  ZBRADJ ZBRCVRT    HMIN    HADJ   HCVRT    HDRY    HWET  BELADJ BELCVRT"});
        t += fmt("    %lf    %lf    %lf    %d    %lf    %lf    %lf    %lf    %lf    %lf    %lf    %lf    %lf\n",
                 c.DX, c.DY, 1.00, 0, c.ZBRADJ, c.ZBRCVRT, c.HMIN, c.HADJ, c.HCVRT, c.HDRY, c.HWET,
                 c.BELADJ, c.BELCVRT);
        cards.push_back({"C11", "C11A", false, t});
    }
    {   // C14 (1445-1466)
        std::string t = lines({
            "C14 TIDAL & ATMOSPHERIC FORCING, GROUND WATER AND SUBGRID CHANNEL PARAMETERS", "*",
            "*   MTIDE:     NUMBER OF PERIOD (TIDAL) FORCING CONSTITUENTS",
            "*   NWSER:     NUMBER OF WIND TIME SERIES (0 SETS WIND TO ZERO)",
            "*   NASER:     NUMBER OF ATMOSPHERIC CONDITION TIME SERIES (0 SETS ALL  ZERO)",
            "*   ISGWI:     1 TO ACTIVATE SOIL MOISTURE BALANCE WITH DRYING AND WETTING",
            "*              2 TO ACTIVATE GROUNDWATER INTERACTION WITH BED AND WATER COL",
            "*  ISCHAN:    >0 ACTIVATE SUBGRID CHANNEL MODEL AND READ MODCHAN.INP",
            "*  ISWAVE:     1-FOR BL IMPACTS (WAVEBL.INP), 2-FOR BL & CURRENT IMPACTS (WAVE.INP)",
            "*              3-FOR INTERNALLY COMPUTED WIND WAVE BOUNDARY LAYER IMPACTS (DS)",
            "* ITIDASM:     1 FOR TIDAL ELEVATION ASSIMILATION (NOT ACTIVE)",
            "*  ISPERC:     1 TO PERCOLATE OR ELIMINATE EXCESS WATER IN DRY CELLS",
            "* ISBODYF:     TO INCLUDE EXTERNAL MODE BODY FORCES FROM FBODY.INP",
            "*              1 FOR UNIFORM OVER DEPTH, 2 FOR SURFACE LAYER ONLY",
            "* ISPNHYDS: 1 FOR QUASI-NONHYDROSTATIC OPTION", "*",
            "C14 MTIDE  NWSER   NASER   ISGWI  ISCHAN  ISWAVE ITIDASM  ISPERC ISBODYF ISPNHYDS"});
        t += fmt("    %d    %d    %d    %d    %d    %d    %d    %d    %d    %d\n", 0, c.NWSER,
                 c.NASER, 0, 0, 0, 0, 0, 0, 0);
        cards.push_back({"C14", "C15", true, t});
    }
    {   // C22 (1493-1513): one concentration series per volume series
        std::string t = lines({
            "C22 SPECIFY NUM OF SEDIMENT AND TOXICS AND NUM OF CONCENTRATION TIME SERIES", "* ",
            "*  NTOX:   NUMBER OF TOXIC CONTAMINANTS (DEFAULT = 1)",
            "*  NSED:   NUMBER OF COHESIVE SEDIMENT SIZE CLASSES (DEFAULT = 1)",
            "*  NSND:   NUMBER OF NON-COHESIVE SEDIMENT SIZE CLASSES (DEFAULT = 1)",
            "*  NCSER1: NUMBER OF SALINITY TIME SERIES",
            "*  NCSER2: NUMBER OF TEMPERATURE TIME SERIES",
            "*  NCSER3: NUMBER OF DYE CONCENTRATION TIME SERIES",
            "*  NCSER4: NUMBER OF SHELLFISH LARVAE CONCENTRATION TIME SERIES",
            "*  NCSER5: NUMBER OF TOXIC CONTAMINANT CONCENTRATION TIME SERIES",
            "*          EACH TIME SERIES MUST HAVE DATA FOR NTOX TOXICICANTS",
            "*  NCSER6: NUMBER OF COHESIVE SEDIMENT CONCENTRATION TIME SERIES",
            "*          EACH TIME SERIES MUST HAVE DATA FOR NSED COHESIVE SEDIMENTS",
            "*  NCSER7: NUMBER OF NON-COHESIVE SEDIMENT CONCENTRATION TIME SERIES",
            "*          EACH TIME SERIES MUST HAVE DATA FOR NSND NON-COHESIVE SEDIMENTS",
            "*  ISSBAL: SET TO 1 FOR SEDIENT MASS BALANCE           ! JOHN & JI, 4/25/97", "*",
            "C22 NTOX    NSED    NSND  NCSER1  NCSER2  NCSER3  NCSER4  NCSER5  NCSER6  NCSER7  ISSBAL"});
        t += fmt("    %d      %d      %d      %d      %d      %d      %d      %d      %d      %d      %d\n",
                 0, 0, 0, salt * c.nqser, temp * c.nqser, dye * c.nqser, 0, 0, 0, 0, 0);
        cards.push_back({"C22", "C23", true, t});
    }
    {   // C23 (1543-1568): the coupling records are the volume sources
        std::string t = lines({
            "C23 VELOCITY, VOLUMN SOURCE / SINK, FLOW CONTROL, AND WITHDRAWAL / RETURN DATA", "*",
            "*  NVBS:   VEL BC(NOT USED)", "*  NUBW : VEL BC(NOT USED)",
            "*  NUBE : VEL BC(NOT USED)", "*  NVBN : VEL BC(NOT USED)",
            "*  NQSIJ : NUMBER OF CONSTANT AND / OR TIME SERIES SPECIFIED SOURCE / SINK",
            "*          LOCATIONS(RIVER INFLOWS, ETC)             .",
            "*  NQJPIJ : NUMBER OF CONSTANT AND / OR TIME SERIES SPECIFIED SOURCE",
            "*          LOCATIONS TREATED AS JETS / PLUMES          .",
            "*  NQSER : NUMBER OF VOLUME SOURCE / SINK TIME SERIES",
            "*  NQCTL : NUMBER OF PRESSURE CONTROLED WITHDRAWAL / RETURN PAIRS",
            "*  NQCTLT : NUMBER OF PRESSURE CONTROLED WITHDRAWAL / RETURN TABLES",
            "*  NQWR : NUMBER OF CONSTANT OR TIME SERIES SPECIFIED WITHDRAWL / RETURN",
            "*          PAIRS",
            "*  NQWRSR : NUMBER OF TIME SERIES SPECIFYING WITHDRAWL, RETURN AND",
            "*          CONCENTRATION RISE SERIES",
            "*  ISDIQ : SET TO 1 TO WRITE DIAGNOSTIC FILE, DIAQ.OUT", "*",
            "C23 NVBS    NUBW    NUBE    NVBN   NQSIJ  NQJPIJ   NQSER   NQCTL  NQCTLT    NQWR  NQWRSR   ISDIQ"});
        t += fmt("    %d      %d      %d      %d      %d      %d      %d      %d      %d      %d     %d      %d\n",
                 0, 0, 0, 0, c.nqsij, 0, c.nqser, 0, 0, 0, 0, 0);
        cards.push_back({"C23", "C24", true, t});
    }
    int nqs = 0;
    {   // C24 (1594-1665): one line per coupling record, edges N, S, E, W,
        // numbered in that order (a record list ends at its first I = 0)
        std::string t = lines({
            "C24 VOLUMETRIC SOURCE / SINK LOCATIONS, MAGNITUDES, AND CONCENTRATION SERIES", "*",
            "*  IQS:      I CELL INDEX OF VOLUME SOURCE / SINK",
            "*  JQS : J CELL INDEX OF VOLUME SOURCE / SINK",
            "*  QSSE : CONSTANT INFLOW / OUTFLOW RATE IN M*m*m / s",
            "*  NQSMUL : MULTIPLIER SWITCH FOR CONSTANT AND TIME SERIES VOL S / S",
            "* = 0  MULT BY 1. FOR NORMAL IN / OUTFLOW(L*L*L / T)",
            "* = 1  MULT BY DY FOR LATERAL IN / OUTFLOW(L*L / T) ON U FACE",
            "* = 2  MULT BY DX FOR LATERAL IN / OUTFLOW(L*L / T) ON V FACE",
            "* = 3  MULT BY DX + DY FOR LATERAL IN / OUTFLOW(L*L / T) ON U&V FACES",
            "*  NQSMFF : IF NON ZERO ACCOUNT FOR VOL S / S MOMENTUM FLUX",
            "* = 1  MOMENTUM FLUX ON NEG U FACE", "* = 2  MOMENTUM FLUX ON NEG V FACE",
            "* = 3  MOMENTUM FLUX ON POS U FACE", "* = 4  MOMENTUM FLUX ON POS V FACE",
            "*  IQSERQ : ID NUMBER OF ASSOCIATED VOLUMN FLOW TIME SERIES",
            "*  ICSER1 : ID NUMBER OF ASSOCIATED SALINITY TIME SERIES",
            "*  ICSER2 : ID NUMBER OF ASSOCIATED TEMPERATURE TIME SERIES",
            "*  ICSER3 : ID NUMBER OF ASSOCIATED DYE CONC TIME SERIES",
            "*  ICSER4 : ID NUMBER OF ASSOCIATED SHELL FISH LARVAE RELEASE TIME SERIES",
            "*  ICSER5 : ID NUMBER OF ASSOCIATED TOXIC CONTAMINANT CONC TIME SERIES",
            "*  ICSER6 : ID NUMBER OF ASSOCIATED COHESIVE SEDIMENT CONC TIME SERIES",
            "*  ICSER7 : ID NUMBER OF ASSOCIATED NON - COHESIVE SED CONC TIME SERIES",
            "*  QSFACTOR : FRACTION OF TIME SERIES FLOW NQSERQ ASSIGNED TO THIS CELL", "*",
            "C24  IQS     JQS     QSSE     NQSMUL  NQSMFF  IQSERQ  ICSER1  ICSER2  ICSER3  ICSER4  ICSER5  ICSER6  ICSER7   QSFACTOR  !ID"});
        static const char *const tag[4] = {"! North_bc", "! South_bc", "! East_bc", "! West_bc"};
        for (int e = 0; e < 4; e++)
            for (const auto &ij : c.bc[e]) {
                if (ij[0] == 0) break;
                nqs = nqs + 1;
                t += fmt("     %d      %d      %lf      %d      %d      %d      %d      %d      %d      %d      %d      %d      %d      %lf      %s\n",
                         ij[0], ij[1], 0.000, 0, 0, nqs, salt * nqs, temp * nqs, dye * nqs, 0, 0, 0,
                         0, 1.000000, tag[e]);
            }
        cards.push_back({"C24", "C25", false, t});
    }
    {   // C25 (1693-1710)
        std::string t = lines({
            "C25 TIME CONSTANT INFLOW CONCENTRATIONS FOR TIME CONSTANT VOLUMETRIC SOURCES", "*",
            "*  SAL: SALT CONCENTRATION CORRESPONDING TO INFLOW ABOVE",
            "*  TEM : TEMPERATURE CORRESPONDING TO INFLOW ABOVE",
            "*  DYE : DYE CONCENTRATION CORRESPONDING TO INFLOW ABOVE",
            "*  SFL : SHELL FISH LARVAE CONCENTRATION CORRESPONDING TO INFLOW ABOVE",
            "*  TOX : NTOX TOXIC CONTAMINANT CONCENTRATIONS CORRESPONDING TO",
            "*       INFLOW ABOVE  WRITTEN AS TOXC(N), N = 1, NTOX A SINGLE DEFAULT",
            "*       VALUE IS REQUIRED EVEN IF TOXIC TRANSPORT IS NOT ACTIVE", "*",
            "C25    SAL       TEM       DYE       SFL  !ID"});
        for (int i = 0; i < nqs; i++)
            t += fmt("     %lf      %lf      %lf      %lf  %s%d\n", 0.00, 0.00, 0.00, 0.00, "!IQSERQ_", i + 1);
        cards.push_back({"C25", "C26", false, t});
    }
    {   // C26 (1743-1760)
        std::string t = lines({
            "C26 TIME CONSTANT INFLOW CONCENTRATIONS FOR TIME CONSTANT VOLUMETRIC SOURCES", "*",
            "*  SED: NSED COHESIVE SEDIMENT CONCENTRATIONS CORRESPONDING TO",
            "*       INFLOW ABOVE  WRITTEN AS SEDC(N), N = 1, NSED.I.E., THE FIRST",
            "*       NSED VALUES ARE COHESIVE A SINGLE DEFAULT VALUE IS REQUIRED",
            "*       EVEN IF COHESIVE SEDIMENT TRANSPORT IS INACTIVE",
            "*  SND : NSND NON - COHESIVE SEDIMENT CONCENTRATIONS CORRESPONDING TO",
            "*       INFLOW ABOVE  WRITTEN AS SND(N), N = 1, NSND.I.E., THE LAST",
            "*       NSND VALUES ARE NON - COHESIVE.A SINGLE DEFAULT VALUE IS",
            "*       REQUIRED EVEN IF NON - COHESIVE SEDIMENT TRANSPORT IS INACTIVE", "*",
            "C26   SED1    SND1"});
        for (int i = 0; i < nqs; i++) t += fmt("     %d      %d\n", 0, 0);
        cards.push_back({"C26", "C27", false, t});
    }
    {   // C71 (1792-1821)
        std::string t = lines({
            "C71 CONTROLS FOR HORIZONTAL PLANE SCALAR FIELD CONTOURING", "*",
            "*  ISSPH:  1 TO WRITE FILE FOR SCALAR FIELD CONTOURING IN HORIZONTAL PLANE",
            "*          2 WRITE ONLY DURING LAST REFERENCE TIME PERIOD",
            "*  NPSPH:    NUMBER OF WRITES PER REFERENCE TIME PERIOD",
            "*  ISRSPH: 1 TO WRITE FILE FOR RESIDUAL SALINITY PLOTTING IN",
            "*            HORIZONTAL",
            "*  ISPHXY: 0 DOES NOT WRITE I,J,X,Y IN ***CNH.OUT AND R***CNH.OUT FILES",
            "*          1 WRITES I,J ONLY IN ***CNH.OUT AND R***CNH.OUT FILES",
            "*          2 WRITES I,J,X,Y  IN ***CNH.OUT AND R***CNH.OUT FILES",
            "*          3 WRITES EFDC_EXPLORER BINARY FORMAT FILES",
            "*  DATA LINE REPEATS 7 TIMES FOR SAL,TEM,DYE,SFL,TOX,SED,SND", "*",
            "C71 ISSPH NPSPH ISRSPH ISPHXY"});
        struct Row { int issph, isphxy; const char *id; };
        const Row rows[] = {{salt, 1, "!SAL"}, {temp, 1, "!TEM"}, {dye, 1, "!DYE"},
                            {0, 3, "!EE WC/Sediment Top Layer Flag"}, {0, 3, "!TOX"},
                            {0, 3, "!SED"}, {0, 3, "!SND"}};
        for (const Row &r : rows) t += fmt("     %d    %d     %d     %d    %s\n", r.issph, 1, 0, r.isphxy, r.id);
        cards.push_back({"C71", "C71A", true, t});
    }
    {   // C72 (1847-1863)
        std::string t = lines({
            "C72 CONTROLS FOR HORIZONTAL SURFACE ELEVATION OR PRESSURE CONTOURING", "*",
            "*  ISPPH:  1 TO WRITE FILE FOR SURFACE ELEVATION OR PRESSURE CONTOURING",
            "* 2 WRITE ONLY DURING LAST REFERENCE TIME PERIOD",
            "*  NPPPH : NUMBER OF WRITES PER REFERENCE TIME PERIOD",
            "*  ISRPPH : 1 TO WRITE FILE FOR RESIDUAL SURFACE ELEVATION  CONTOURNG IN",
            "*            HORIZONTAL PLANE",
            "*  IPPHXY : 0 DOES NOT WRITE I, J, X, Y IN surfplt.out and rsurfplt.out FILES",
            "* 1 WRITES I, J ONLY IN surfplt.out and rsurfplt.out FILES",
            "* 2 WRITES I, J, X, Y  IN surfplt.out and rsurfplt.out FILES",
            "* 3 WRITES EFDC EXPLORER BINARY FORMAT FILES", "*",
            "C72 ISPPH NPPPH ISRPPH IPPHXY"});
        t += fmt("     %d    %d     %d     %d\n", c.ISPPH, c.NPPPH, c.ISRPPH, c.IPPHXY);
        cards.push_back({"C72", "C73", true, t});
    }
    {   // C73 (1889-1905)
        std::string t = lines({
            "C73 CONTROLS FOR HORIZONTAL PLANE VELOCITY VECTOR PLOTTING", "*",
            "*  ISVPH:  1 TO WRITE FILE FOR VELOCITY PLOTTING IN HORIZONTAL PLANE",
            "* 2 WRITE ONLY DURING LAST REFERENCE TIME PERIOD",
            "*  NPVPH : NUMBER OF WRITES PER REFERENCE TIME PERIOD",
            "*  ISRVPH : 1 TO WRITE FILE FOR RESIDUAL VELOCITY PLOTTIN IN",
            "*            HORIZONTAL PLANE",
            "*  IVPHXY : 0 DOES NOT WRITE I, J, X, Y IN velplth.out and rvelplth.out FILES",
            "* 1 WRITES I, J ONLY IN velplth.out and rvelplth.out FILES",
            "* 2 WRITES I, J, X, Y  IN velplth.out and rvelplth.out FILES",
            "* 3 WRITES EFDC EXPLORER BINARY FORMAT FILES", "*",
            "C73 ISVPH NPVPH ISRVPH IVPHXY"});
        t += fmt("     %d     %d     %d      %d\n", c.ISVPH, c.NPVPH, c.ISRVPH, c.IVPHXY);
        cards.push_back({"C73", "C74", false, t});
    }
    {   // C84 (1931-1951)
        std::string t = lines({
            "C84 CONTROLS FOR WRITING TO TIME SERIES FILES", "*",
            "*  ISTMSR:  1 OR 2 TO WRITE TIME SERIES OF SURF ELEV, VELOCITY, NET",
            "*           INTERNAL AND EXTERNAL MODE VOLUME SOURCE - SINKS, AND",
            "*           CONCENTRATION VARIABLES, 2 APPENDS EXISTING TIME SERIES FILES",
            "*  MLTMSR : NUMBER HORIZONTAL LOCATIONS TO WRITE TIME SERIES OF SURF ELEV,",
            "*VELOCITY, AND CONCENTRATION VARIABLES",
            "*  NBTMSR : TIME STEP TO BEGIN WRITING TO TIME SERIES FILES(Inactive)",
            "*  NSTMSR : TIME STEP TO STOP WRITING TO TIME SERIES FILES(Inactive)",
            "*  NWTMSR : NUMBER OF TIME STEPS TO SKIP BETWEEN OUTPUT",
            "*  NTSSTSP : NUMBER OF TIME SERIES START - STOP SCENARIOS, 1 OR GREATER",
            "*  TCTMSR : UNIT CONVERSION FOR TIME SERIES TIME.FOR SECONDS, MINUTES,",
            "*HOURS, DAYS USE 1.0, 60.0, 3600.0, 86400.0 RESPECTIVELY", "*", "*",
            "C84 ISTMSR MLTMSR  NBTMSR  NSTMSR  NWTMSR NTSSTSP TCTMSR"});
        t += fmt("     %d    %d     %d    %d    %d    %d     %lf\n", c.ISTMSR, c.MLTMSR, c.NBTMSR,
                 c.NSTMSR, c.NWTMSR, c.NTSSTSP, c.TCTMSR);
        cards.push_back({"C84", "C85", true, t});
    }
    {   // C87 (1977-2012): one line per EFDC time-series location
        std::string t = lines({
            "C87 CONTROLS FOR WRITING TO TIME SERIES FILES", "*",
            "*  ILTS:    I CELL INDEX", "*  JLTS : J CELL INDEX",
            "*  NTSSSS : WRITE SCENARIO FOR THIS LOCATION",
            "*  MTSP : 1 FOR TIME SERIES OF SURFACE ELEVATION",
            "*  MTSC : 1 FOR TIME SERIES OF TRANSPORTED CONCENTRATION VARIABLES",
            "*  MTSA : 1 FOR TIME SERIES OF EDDY VISCOSITY AND DIFFUSIVITY",
            "*  MTSUE : 1 FOR TIME SERIES OF EXTERNAL MODE HORIZONTAL VELOCITY",
            "*  MTSUT : 1 FOR TIME SERIES OF EXTERNAL MODE HORIZONTAL TRANSPORT",
            "*  MTSU : 1 FOR TIME SERIES OF HORIZONTAL VELOCITY IN EVERY LAYER",
            "*  MTSQE : 1 FOR TIME SERIES OF NET EXTERNAL MODE VOLUME SOURCE / SINK",
            "*  MTSQ : 1 FOR TIME SERIES OF NET EXTERNAL MODE VOLUME SOURCE / SINK",
            "*  CLTS : LOCATION AS A CHARACTER VARIALBLE", "*",
            "C87 ILTS JLTS NTSSSS MTSP MTSC MTSA MTSUE MTSUT MTSU MTSQE MTSQ CLTS"});
        for (size_t i = 0; i < c.LTS.size(); i++)
            t += fmt("     %d    %d    %d    %d    %d    %d    %d    %d    %d    %d    %d  %s%d%s\n",
                     c.LTS[i][0], c.LTS[i][1], c.NTSSSS, c.MTSP, c.MTSC, c.MTSA, c.MTSUE, c.MTSUT,
                     c.MTSU, 0, 0, "'P_", (int)i, "'");
        cards.push_back({"C87", "C88", true, t});
    }

    for (const Card &k : cards) {
        const Token next = find_token(s, k.next);
        const Token here = find_token(s, k.name);
        if (!next.found || !here.found) {
            if (err) *err = std::string("EFDC.INP has no card ") + (here.found ? k.next : k.name);
            return false;
        }
        // kept tail: past the successor's token, its next character and the
        // whitespace after that
        size_t tail = std::min(next.end + 1, s.size());
        while (tail < s.size() && isspace((unsigned char)s[tail])) tail++;
        std::string out = s.substr(0, here.prev_end);
        if (k.lead) out += '\n';
        out += k.text;
        out += kDashes;
        out += '\n';
        out += k.next;
        out += ' ';
        out.append(s, tail, std::string::npos);
        s.swap(out);
    }
    // written beside and renamed over the file, so a failed write leaves the
    // previous file whole
    const std::string tmp = path + ".tmp";
    FILE *fp = fopen(tmp.c_str(), "wb");
    if (!fp) {
        if (err) *err = "cannot write " + tmp;
        return false;
    }
    bool ok = fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (fclose(fp) == 0) && ok;
    if (ok && rename(tmp.c_str(), path.c_str()) != 0) ok = false;
    if (!ok) {
        remove(tmp.c_str());
        if (err) *err = "cannot write " + path;
    }
    return ok;
}

}  // namespace sw2dio
