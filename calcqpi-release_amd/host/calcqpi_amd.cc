// calcqpi_amd -- calcqpi-compatible driver for the continuum-LDOS (wannier)
// mode on one MI355X, on top of the C ABI of libcqpi_amd.so.
//
//   calcqpi_amd <inputfile>             run and write the qpifile (idl map)
//   calcqpi_amd --print-config <input>  parse only, print the resolved job as JSON
//
// The input file and the output map follow the reference driver:
//   * key=value; lines, '#' comments, (a,b,...) lists, "string" values,
//     a warning for every unused key (loadconfig, include/parser.h:24-199);
//   * defaults and keys of main(), src/calcqpi.cc:133-273, for output=wannier:
//     tbfile, qpifile, lattice, oversamp, window, layers,
//     energies, eta, kpoints, fermi, scattering, phase, spin, scmodel,
//     magscat, prearr, idlorbitalfile, logfile;
//   * the orbital tiles come from the idl orbital file, duplicated for spin
//     or padded with zero tiles as in src/calcqpi.cc:483-518;
//   * energies of the layers are lb + (hb-lb) k/(L-1) (idl::getbias,
//     src/idl.cc:969-973); the map is W x W x L with W = oversamp*lattice,
//     header as in src/calcqpi.cc:640,655 and data written as float32, x
//     fastest (idl::write, src/idl.cc:78-95).
//   * output=josephson (scmodel models only) with deltat and etat, the tip
//     factor of src/calcqpi.cc:705-709, header name of src/calcqpi.cc:643-645;
//   * threshold=t builds the threshold list (src/calcqpi.cc:620-634) and
//     computes the wannier map from it (src/calcqpi.cc:699-702);
//   * output=spf|uspf: spectral-function maps n x n (spf sets kpoints = n,
//     src/calcqpi.cc:333-334; uspf takes pos[i] and prearr, :464-482),
//     header names of :647-652; output=nomode makes no map;
//   * dosfile, doslayers, dosenergies: the density of states as text lines
//     "energy dos" (src/calcqpi.cc:336-391).
//   * orbitals=(s,px,...) with radius, angle, anglearr, zheight, pos[i] and
//     prearr: analytic orbitals sampled on the tile grid (src/calcqpi.cc:
//     519-575); window=0 sizes the window from the radius; wffile writes
//     the tiles (:602-618);
//   * bsfile, bslattice, bslayers, bsenergies, bsoversamp: the band-structure
//     histogram as an idl file (src/calcqpi.cc:308-326);
//   * green=surface|bulk with stbfile and epserr: the surface or bulk
//     Green's function of a layered model for every output (src/calcqpi.cc:
//     203-208, :293-307, :367-375); the principal-layer warning as there.
// Extension key: precision=fp64|fp32|mixed; (contraction precision, default fp64).
//   * orbitalfiles=("a.xsf",...) with zheight: Wannier90/XSF orbital files
//     (src/calcqpi.cc:576-601).
//   * spinarr: the threshold list's spin labels (src/calcqpi.cc:626-628).
// Other output modes are reported as unsupported.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cqpi_amd.h"

using std::string;
using std::vector;

namespace {

// ---------------------------------------------------------------------------
// input file: key=value; (include/parser.h:29-54 semantics)
// ---------------------------------------------------------------------------
struct Config {
    vector<string> name, field;
    vector<int> used;
    bool ok = false;

    explicit Config(const char *path)
    {
        std::ifstream is(path);
        if (!is) return;
        ok = true;
        string line;
        size_t lineno = 0;
        while (std::getline(is, line)) {
            if (line.empty()) continue;
            lineno++;
            size_t hash = line.find('#');
            if (hash != string::npos) line.erase(hash);
            size_t eq = line.find('='), semi = line.find(';');
            if (eq != string::npos && semi != string::npos) {
                name.push_back(line.substr(0, eq));
                field.push_back(line.substr(eq + 1, semi - eq - 1));
            } else if (eq != string::npos || semi != string::npos) {
                std::cerr << "Incorrect syntax, '=' or ';' missing.\nLine " << lineno << ": " << line << "\n";
            }
        }
        used.assign(name.size(), 0);
    }
    bool probe(const string &key)
    {
        for (size_t i = 0; i < name.size(); i++)
            if (name[i] == key) {
                used[i]++;
                return true;
            }
        return false;
    }
    string get(const string &key)
    {
        string v;
        for (size_t i = 0; i < name.size(); i++)
            if (name[i] == key) {
                used[i]++;
                v = field[i];
            }
        if (v.empty()) return v;
        size_t a = v.find_first_not_of(" \t"), b = v.find_last_not_of(" \t");
        if (a == string::npos) return "";
        return v.substr(a, b - a + 1);
    }
    static string unquote(const string &v)
    {
        size_t a = v.find_first_of('"'), b = v.find_last_of('"');
        if (a != string::npos && b != string::npos && b > a) return v.substr(a + 1, b - a - 1);
        return "";
    }
    string str(const string &key) { return unquote(get(key)); }
    double num(const string &key)
    {
        string v = get(key);
        return v.empty() ? NAN : std::atof(v.c_str());
    }
    int integer(const string &key)
    {
        string v = get(key);
        return v.empty() ? 0 : std::atoi(v.c_str());
    }
    vector<string> list(const string &key)
    {
        string v = get(key);
        vector<string> out;
        size_t first = v.find('('), last = v.rfind(')');
        if (first == string::npos || last == string::npos || last < first) return out;
        string inner = v.substr(first + 1, last - first - 1);
        std::stringstream ss(inner);
        string item;
        while (std::getline(ss, item, ',')) {
            size_t a = item.find_first_not_of(" \t"), b = item.find_last_not_of(" \t");
            out.push_back(a == string::npos ? string() : item.substr(a, b - a + 1));
        }
        return out;
    }
    vector<double> vec(const string &key)
    {
        vector<double> out;
        for (auto &s : list(key)) out.push_back(std::atof(s.c_str()));
        return out;
    }
    void warn_unused(std::ostream &os)
    {
        for (size_t i = 0; i < name.size(); i++)
            if (!used[i]) os << "Unused line " << i << ": " << name[i] << "\n";
    }
};

// ---------------------------------------------------------------------------
// idl files (src/idl.cc:50-95)
// ---------------------------------------------------------------------------
// Stream buffer that writes through to another one and starts every line
// with the local time "YYYY-MM-DD.HH:MM:SS " (the reference's log file,
// AddTimeStamp, include/addtimestamp.h:25-84).
class StampedLines : public std::streambuf {
    std::streambuf *sink_;
    bool at_line_start_ = true;

  protected:
    int overflow(int c) override
    {
        if (c == traits_type::eof()) return sink_->pubsync() == -1 ? c : traits_type::not_eof(c);
        if (at_line_start_) {
            char buf[80];
            const time_t now = time(nullptr);
            struct tm t;
            localtime_r(&now, &t);
            const size_t len = strftime(buf, sizeof buf, "%Y-%m-%d.%X ", &t);
            if (sink_->sputn(buf, (std::streamsize)len) != (std::streamsize)len) return traits_type::eof();
        }
        at_line_start_ = c == '\n';
        const int r = sink_->sputc((char)c);
        if (at_line_start_) sink_->pubsync();
        return r;
    }
    int sync() override { return sink_->pubsync(); }

  public:
    explicit StampedLines(std::streambuf *sink) : sink_(sink) {}
};

// Wall and process-CPU time per named section, sections of one name summed,
// printed as the reference's timing summary (recordtimings,
// include/recordtimings.h:26-113).
class SectionTimes {
    struct Entry {
        string name;
        double real, cpu;
    };
    vector<Entry> entries_;
    double last_real_, last_cpu_;
    static double now(clockid_t id)
    {
        timespec t;
        clock_gettime(id, &t);
        return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
    }
    static string span(double sec)
    {
        const long long s = (long long)sec;
        char buf[64];
        snprintf(buf, sizeof buf, "%2lldd, %2lldh, %2lldm, %2llds%7lld\u03bcs", s / 86400, (s % 86400) / 3600,
                 (s % 3600) / 60, s % 60, (long long)((sec - (double)s) * 1e6));
        return buf;
    }

  public:
    SectionTimes() : last_real_(now(CLOCK_REALTIME)), last_cpu_(now(CLOCK_PROCESS_CPUTIME_ID)) {}
    void mark(const string &name)
    {
        const double r = now(CLOCK_REALTIME), c = now(CLOCK_PROCESS_CPUTIME_ID);
        auto it = std::find_if(entries_.begin(), entries_.end(), [&](const Entry &e) { return e.name == name; });
        if (it == entries_.end()) entries_.push_back({name, r - last_real_, c - last_cpu_});
        else {
            it->real += r - last_real_;
            it->cpu += c - last_cpu_;
        }
        last_real_ = r;
        last_cpu_ = c;
    }
    void print(std::ostream &os) const
    {
        double tr = 0.0, tc = 0.0;
        for (const Entry &e : entries_) {
            tr += e.real;
            tc += e.cpu;
        }
        const string rule(114, '-');
        os << "Summary of timings: \n"
           << "     Section                            CPU time                               Real time\n"
           << " " << rule << "\n";
        for (const Entry &e : entries_) {
            char line[256];
            snprintf(line, sizeof line, "     %-33s %s (%5.2g%%)   %s (%5.2g%%)", e.name.c_str(), span(e.cpu).c_str(),
                     tc > 0 ? 100.0 * e.cpu / tc : 0.0, span(e.real).c_str(), tr > 0 ? 100.0 * e.real / tr : 0.0);
            os << line << "\n";
        }
        os << " " << rule << "\n"
           << "     Total time elapsed                " << span(tc) << "            " << span(tr) << "\n";
    }
};

struct IdlMap {
    string name, date;
    size_t sx = 0, sy = 0, sz = 0;
    double hdr[7] = {0, 0, 0, 0, 0, 0, 0};  // scan x,y; offset x,y; bias lo,hi; setpoint
    vector<double> data;                   // x + y*sx + z*sx*sy
};

bool read_idl(const string &path, IdlMap &m)
{
    std::ifstream is(path, std::ios::binary);
    if (!is) return false;
    std::getline(is, m.name);
    std::getline(is, m.date);
    is >> m.sx >> m.sy >> m.sz;
    for (double &v : m.hdr) is >> v;
    if (!is) return false;
    is.ignore(1, '\n');
    size_t tot = m.sx * m.sy * m.sz;
    vector<float> raw(tot);
    is.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)(tot * sizeof(float)));
    if (!is) return false;
    m.data.assign(raw.begin(), raw.end());
    return true;
}

bool write_idl(const string &path, const IdlMap &m)
{
    std::ofstream os(path, std::ios::binary);
    if (!os) return false;
    os << m.name << "\n" << m.date << "\n" << m.sx << "\n" << m.sy << "\n" << m.sz << "\n";
    for (double v : m.hdr) os << v << "\n";
    vector<float> raw(m.data.begin(), m.data.end());
    os.write(reinterpret_cast<const char *>(raw.data()), (std::streamsize)(raw.size() * sizeof(float)));
    return (bool)os;
}

int check(int rc, const char *what)
{
    if (rc) std::cerr << what << " failed: " << cqpi_last_error() << " (code " << rc << ")\n";
    return rc;
}

string json_escape(const string &s)
{
    string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

}  // namespace

int main(int argc, char **argv)
{
    bool print_only = false;
    int argi = 1;
    if (argc > 2 && string(argv[1]) == "--print-config") {
        print_only = true;
        argi = 2;
    }
    if (argi >= argc) {
        std::cerr << "Syntax: calcqpi_amd [--print-config] configfile\n"
                  << "Continuum-LDOS (output=wannier|josephson, green=normal) on one MI355X; keys as calcqpi,\n"
                  << "plus precision=fp64|fp32|mixed; for the contraction.\n";
        return 1;
    }
    Config lf(argv[argi]);
    if (!lf.ok) {
        std::cerr << "Error while opening file " << argv[argi] << " for reading.\n";
        return 1;
    }
    std::ofstream logfile;
    std::ostream *log = print_only ? &std::cerr : &std::cout;  // --print-config keeps stdout for the JSON
    // logfile: every line time-stamped, and error messages go there too
    // (src/calcqpi.cc:162-170)
    std::unique_ptr<StampedLines> stamped;
    std::unique_ptr<std::ostream> stamped_log;
    struct CerrRestore {
        std::streambuf *saved = nullptr;
        ~CerrRestore()
        {
            if (saved) std::cerr.rdbuf(saved);
        }
    } cerr_restore;
    if (lf.probe("logfile")) {
        logfile.open(lf.str("logfile"), std::ios::app);
        if (logfile.is_open()) {
            stamped.reset(new StampedLines(logfile.rdbuf()));
            stamped_log.reset(new std::ostream(stamped.get()));
            log = stamped_log.get();
            if (!print_only) cerr_restore.saved = std::cerr.rdbuf(stamped.get());
        }
    }
    *log << "New calculation started, using configuration from " << argv[argi] << "\n";
    // defaults: src/calcqpi.cc:133-137
    long n = 201, layers = 21, oversamp = 4, window = 2, kpoints = 0;
    bool spin = false, scmodel = false;
    double lenergy = -1.0, henergy = 1.0, eta = 0.005, efermi = 0.0, magscat = 0.0;
    double threshold = NAN, deltat = 0.001, etat = NAN;
    bool josephson = false;
    int spectral = 0;     // 1: spf, 2: uspf
    bool nomap = false;   // output=nomode
    std::complex<double> scatphase(1.0, 0.0);
    vector<std::complex<double>> scattering;
    string tbfile, qpifile, precision = "fp64";

    if (lf.probe("tbfile")) tbfile = lf.str("tbfile");
    else {
        std::cerr << "No tightbinding model specified ... exiting.\n";
        return 1;
    }
    if (lf.probe("fermi")) efermi = lf.num("fermi");
    // Green's function: src/calcqpi.cc:129, :203-208 (anything else is normal)
    int green = CQPI_GREEN_NORMAL;
    double epserr = 1.0e-5;
    string stbfile;
    if (lf.probe("stbfile")) stbfile = lf.str("stbfile");
    if (lf.probe("epserr")) epserr = lf.num("epserr");
    if (lf.probe("green")) {
        if (lf.get("green") == "surface") green = CQPI_GREEN_SURFACE;
        else if (lf.get("green") == "bulk") green = CQPI_GREEN_BULK;
    }
    if (lf.probe("spin")) spin = lf.get("spin") == "true";
    if (lf.probe("scmodel")) scmodel = lf.get("scmodel") == "true";
    if (lf.probe("output")) {
        const string mode = lf.get("output");
        if (mode == "josephson") josephson = true;
        else if (mode == "spf") spectral = 1;
        else if (mode == "uspf") spectral = 2;
        else if (mode == "nomode") nomap = true;
        else if (mode != "wannier") {
            std::cerr << "output=" << mode << " is not supported by this build (wannier, josephson, spf, uspf, "
                      << "nomode).\n";
            return 1;
        }
    }
    if (lf.probe("threshold")) threshold = lf.num("threshold");
    // spin label per orbital for the threshold list (src/calcqpi.cc:626-628;
    // read as the reference's getintvector, atoi per entry)
    bool have_spinarr = false;
    vector<int> spinarr;
    if (lf.probe("spinarr")) {
        have_spinarr = true;
        for (auto &v : lf.list("spinarr")) spinarr.push_back(std::atoi(v.c_str()));
    }
    if (lf.probe("deltat")) deltat = lf.num("deltat");
    if (lf.probe("etat")) etat = lf.num("etat");
    if (josephson && !scmodel) {
        // src/calcqpi.cc:458-461
        std::cerr << "Error: Calculation in josephson-mode with scmodel=false makes no sense - exiting.\n";
        return 1;
    }
    if (lf.probe("qpifile")) qpifile = lf.str("qpifile");
    if (lf.probe("lattice")) n = lf.integer("lattice");
    if (lf.probe("oversamp")) oversamp = lf.integer("oversamp");
    if (lf.probe("window")) window = lf.integer("window");
    if (lf.probe("layers")) layers = lf.integer("layers");
    if (lf.probe("energies")) {
        vector<double> en = lf.vec("energies");
        if (en.size() >= 2) {
            lenergy = en[0];
            henergy = en[1];
        }
    }
    if (lf.probe("eta")) eta = lf.num("eta");
    if (std::isnan(etat)) etat = eta;  // src/calcqpi.cc:249-250
    if (lf.probe("kpoints")) kpoints = lf.integer("kpoints");
    if (!kpoints) kpoints = n;
    if (spectral == 1) kpoints = n;  // src/calcqpi.cc:333-334
    // density of states: src/calcqpi.cc:336-391
    string dosfile;
    long doslayers = 101;
    double doslenergy = -0.1, doshenergy = 0.1;
    if (lf.probe("dosfile")) {
        dosfile = lf.str("dosfile");
        if (lf.probe("doslayers")) doslayers = lf.integer("doslayers");
        if (lf.probe("dosenergies")) {
            vector<double> en = lf.vec("dosenergies");
            if (en.size() >= 2) {
                doslenergy = en[0];
                doshenergy = en[1];
            }
        }
    }
    if (lf.probe("phase")) {
        vector<double> ph = lf.vec("phase");
        if (ph.size() >= 2) scatphase = std::complex<double>(ph[0], ph[1]);
    }
    if (lf.probe("scattering"))
        for (double v : lf.vec("scattering")) scattering.push_back(v * scatphase);
    if (lf.probe("magscat")) magscat = lf.num("magscat");
    if (lf.probe("precision")) precision = lf.get("precision");
    if (precision != "fp64" && precision != "fp32" && precision != "mixed") {
        std::cerr << "precision must be fp64, fp32 or mixed.\n";
        return 1;
    }

    // model
    cqpi_hopping *hops = nullptr;
    size_t nhops = 0, bands = 0;
    char tbname[1024];
    if (check(cqpi_read_hr(tbfile.c_str(), scmodel, &hops, &nhops, &bands, tbname, sizeof tbname), "reading the TB model")) {
        std::cerr << "Error loading TB model from " << tbfile << ", exiting.\n";
        return 1;
    }
    // surface layer model and the principal-layer check (src/calcqpi.cc:293-307)
    cqpi_hopping *shops = nullptr;
    size_t nshops = 0;
    auto principal = [](const cqpi_hopping *h, size_t nh) {
        for (size_t i = 0; i < nh; i++)
            if (std::fabs(h[i].rz) > 1 && (h[i].tre != 0.0 || h[i].tim != 0.0)) return false;
        return true;
    };
    if (!stbfile.empty()) {
        size_t sbands = 0;
        char sname[1024];
        *log << "Loading surface TB model " << stbfile << "\n";
        if (check(cqpi_read_hr(stbfile.c_str(), scmodel, &shops, &nshops, &sbands, sname, sizeof sname),
                  "reading the surface TB model") ||
            sbands != bands) {
            std::cerr << "Error loading surface TB model from " << stbfile << ", exiting.\n";
            cqpi_free(hops);
            cqpi_free(shops);
            return 1;
        }
    }
    if (green != CQPI_GREEN_NORMAL) {
        if (!principal(hops, nhops))
            *log << "Warning: Tight-binding model in file " << tbfile << " is not in principal layer form.\n";
        if (shops && !principal(shops, nshops))
            *log << "Warning: Surface tight-binding model in file " << stbfile
                 << " is not in principal layer form.\n";
    }
    size_t maxbands = bands;
    if (scmodel) {
        maxbands = bands >> 1;
        if (scattering.size() == maxbands)
            for (size_t i = 0; i < maxbands; i++) scattering.push_back((-1.0 + 2.0 * magscat) * scattering[i]);
    }
    // V = diag(vscat): src/qpi.h:107-130, src/calcqpi.cc:393-395
    vector<std::complex<double>> vscat(bands, scatphase);
    if (!scattering.empty())
        for (size_t i = 0; i < bands; i++) vscat[i] = i < scattering.size() ? scattering[i] : 0.0;

    // unfolded spectral function: orbital positions and prefactors (src/calcqpi.cc:464-482)
    vector<double> posxy(2 * maxbands, 0.0), sprearr;
    if (spectral == 2) {
        for (size_t i = 0; i < maxbands; i++) {
            const string key = "pos[" + std::to_string(i) + "]";
            if (lf.probe(key)) {
                vector<double> p = lf.vec(key);
                for (size_t c = 0; c < 2 && c < p.size(); c++) posxy[2 * i + c] = p[c];
            } else
                *log << "No position found for band " << i << ", using (0,0,0).\n";
        }
        if (lf.probe("prearr")) sprearr = lf.vec("prearr");
        sprearr.resize(maxbands, 1.0);
    }
    // Wannier tiles (src/calcqpi.cc:483-617): from the idl orbital file, or
    // analytic orbitals sampled on the tile grid (orbitals=, radius, angle,
    // anglearr, zheight, pos[i], prearr; window = 0 sizes the window)
    // (read in every output mode, as the reference does; the spectral maps and
    // the DOS then leave them unused)
    vector<vector<double>> wfs;
    if (lf.probe("idlorbitalfile")) {
        const size_t T = (size_t)(2 * window + 1) * oversamp;
        string orbfile = lf.str("idlorbitalfile");
        *log << "Using IDL orbital file for vacuum overlap.\nLoading wave function file " << orbfile << ".\n";
        IdlMap orb;
        if (!read_idl(orbfile, orb)) {
            std::cerr << "Error loading orbitals from file " << orbfile << ", exiting.\n";
            return 1;
        }
        if (orb.sx != orb.sy || orb.sx != T) {
            std::cerr << "Orbital functions in " << orbfile << " have unsuitable dimensions.\n";
            return 1;
        }
        vector<double> prearr;
        if (lf.probe("prearr")) prearr = lf.vec("prearr");
        prearr.resize(std::max(orb.sz, maxbands), 1.0);
        for (size_t k = 0; k < orb.sz; k++) {
            vector<double> t(T * T);
            for (size_t i = 0; i < T; i++)
                for (size_t j = 0; j < T; j++) t[i * T + j] = prearr[k] * orb.data[i + j * T + k * T * T];
            wfs.push_back(t);
        }
        if ((maxbands >> 1) == orb.sz && spin) {
            *log << "Found half as many orbitals as needed in spin-polarized calculation, duplicating wave "
                    "function set.\n";
            for (size_t k = 0; k < orb.sz; k++) wfs.push_back(wfs[k]);
        } else if (maxbands > orb.sz) {
            *log << "Warning: Fewer number of orbitals in " << orbfile << " than there are bands, padding with zeros.\n";
        }
    } else if (lf.probe("orbitals")) {
        vector<string> names = lf.list("orbitals");
        double radius = 0.5, angle = 0.0, zheight = 5.0;
        if (lf.probe("radius")) radius = lf.num("radius");
        if (lf.probe("angle")) angle = lf.num("angle");
        if (lf.probe("zheight")) zheight = lf.num("zheight");
        *log << "Using gaussian orbitals for vacuum overlap.\n    Parameters:\n             Radius:         " << radius
             << "\n             Rotation angle: " << angle << "\n             Height:         " << zheight << "\n";
        vector<double> anglearr;
        if (lf.probe("anglearr")) anglearr = lf.vec("anglearr");
        anglearr.resize(names.size(), 0.0);
        vector<double> prearr;
        if (lf.probe("prearr")) prearr = lf.vec("prearr");
        prearr.resize(names.size(), 1.0);
        struct Orb {
            string name;
            double height, x, y, theta, pre;
        };
        vector<Orb> orbs;
        for (size_t i = 0; i < names.size(); i++) {
            const string key = "pos[" + std::to_string(i) + "]";
            const double th = (angle + anglearr[i]) * M_PI / 180.0;
            if (lf.probe(key)) {
                vector<double> p = lf.vec(key);
                p.resize(3, 0.0);
                orbs.push_back({names[i], zheight - p[2], p[0], p[1], th, prearr[i]});
            } else {
                orbs.push_back({names[i], zheight, 0.0, 0.0, th, prearr[i]});
            }
        }
        if ((maxbands >> 1) == names.size() && spin) {
            *log << "Found half as many orbitals as needed in spin-polarized calculation, duplicating wave "
                    "function set.\n";
            for (size_t k = 0; k < names.size(); k++) orbs.push_back(orbs[k]);
        } else if (names.size() < maxbands) {
            *log << "Warning: Fewer number of orbitals specified than there are bands, padding with zeros.\n";
            for (size_t i = names.size(); i < maxbands; i++) orbs.push_back({"z", zheight, 0.0, 0.0, 0.0, 0.0});
        }
        *log << "Parameters for Wannier functions:\n";
        for (size_t i = 0; i < orbs.size(); i++)
            *log << "Orbital " << i << ": " << orbs[i].name << "@(" << orbs[i].x << "," << orbs[i].y << "), height "
                 << orbs[i].height << ", theta " << orbs[i].theta * 180.0 / M_PI << "deg \n";
        if (window) {
            const double val = radius * std::sqrt(std::log(100.0) + std::fabs(zheight));
            *log << "Recommended window size for r=" << radius << ": " << std::ceil(val * val - zheight * zheight)
                 << "\nchosen window size: " << window << "\n";
        } else {
            window = (long)std::ceil(cqpi_orbital_range(radius, 0.01));
            *log << "Using recommended window size for r=" << radius << ": " << window << "\n";
        }
        const size_t T = (size_t)(2 * window + 1) * oversamp;
        *log << "Precalculating wave functions.\n";
        for (const Orb &o : orbs) {
            vector<double> t(T * T);
            if (cqpi_orbital_tile(o.name.c_str(), radius, o.height, o.x, o.y, o.theta, o.pre, (size_t)window,
                                  (size_t)oversamp, t.data())) {
                std::cerr << "Unknown orbital type " << o.name << ", exiting.\n";
                return 1;
            }
            wfs.push_back(t);
        }
        *log << "Done.\n";
        // largest boundary value relative to each tile's maximum (checkwavefunctions, src/calcqpi.cc:118-127)
        double worst = 0.0;
        for (const vector<double> &t : wfs) {
            double mx = 0.0, mb = 0.0;
            for (size_t i = 0; i < T; i++)
                for (size_t j = 0; j < T; j++) {
                    const double v = std::fabs(t[i * T + j]);
                    mx = std::max(mx, v);
                    if (i == 0 || j == 0 || i == T - 1 || j == T - 1) mb = std::max(mb, v);
                }
            if (mx > 0.0) worst = std::max(worst, mb / mx);
        }
        *log << "Maximum rel. value at boundary: " << worst << "\n";
        if (worst > 0.05) *log << "Warning: boundary too close, use larger window size!\n";
    } else if (lf.probe("orbitalfiles")) {
        // Wannier90/XSF orbital files: src/calcqpi.cc:576-601
        double zheight = 5.0;
        if (lf.probe("zheight")) zheight = lf.num("zheight");
        *log << "Using Wannier90/XSF-orbitals for vacuum overlap.\n    Parameters:\n             zheight:         "
             << zheight << "\n";
        vector<string> files = lf.list("orbitalfiles");
        for (string &f : files)
            if (f.size() >= 2 && f.front() == '"' && f.back() == '"') f = f.substr(1, f.size() - 2);
        vector<double> prearr;
        if (lf.probe("prearr")) prearr = lf.vec("prearr");
        prearr.resize(files.size(), 1.0);
        const size_t T = (size_t)(2 * window + 1) * oversamp;
        for (size_t i = 0; i < files.size(); i++) {
            *log << "Loading wave function file " << files[i] << ".\n";
            vector<double> t(T * T);
            if (cqpi_xsf_tile(files[i].c_str(), zheight, (size_t)window, (size_t)oversamp, prearr[i], t.data())) {
                std::cerr << "Error loading wave function file " << files[i] << "\n";
                return 1;
            }
            wfs.push_back(t);
        }
        if ((maxbands >> 1) == files.size() && spin) {
            *log << "Found half as many orbitals as needed in spin-polarized calculation, duplicating wave "
                    "function set.\n";
            for (size_t k = 0; k < files.size(); k++) wfs.push_back(wfs[k]);
        } else if (files.size() < maxbands) {
            *log << "Fewer number of orbital files specified than there are bands, padding with zeros.\n";
        }
    } else if (!(spectral || nomap)) {
        std::cerr << "The Wannier tiles need an idlorbitalfile, analytic orbitals (orbitals=) or orbital files "
                     "(orbitalfiles=).\n";
        return 1;
    }
    const size_t T = (size_t)(2 * window + 1) * oversamp;
    vector<double> tiles(maxbands * T * T, 0.0);
    for (size_t o = 0; o < maxbands && o < wfs.size(); o++)
        std::copy(wfs[o].begin(), wfs[o].end(), tiles.begin() + o * T * T);
    if (lf.probe("wffile") && !wfs.empty() && !print_only) {
        // the tiles as an idl file (src/calcqpi.cc:602-618)
        IdlMap wf;
        wf.sx = wf.sy = T;
        wf.sz = wfs.size();
        wf.data.assign(T * T * wfs.size(), 0.0);
        for (size_t l = 0; l < wfs.size(); l++)
            for (size_t i = 0; i < T; i++)
                for (size_t j = 0; j < T; j++) wf.data[i + j * T + l * T * T] = wfs[l][i * T + j];
        std::ostringstream nm;
        nm << "Wannier function file for " << tbfile << " with window=" << window << " and oversamp=" << oversamp;
        wf.name = nm.str();
        wf.hdr[0] = wf.hdr[1] = 1.0;  // idl(x, y, z) defaults: scan range 1 x 1, offset and bias 0
        const time_t now = time(nullptr);
        wf.date = ctime(&now);
        if (!wf.date.empty() && wf.date.back() == '\n') wf.date.pop_back();
        const string wfname = lf.str("wffile");
        *log << "Writing wave function file to " << wfname << "\n";
        if (!write_idl(wfname, wf)) {
            std::cerr << "Error writing " << wfname << "\n";
            return 1;
        }
    }
    vector<double> omegas(layers);
    for (long k = 0; k < layers; k++)
        omegas[k] = layers == 1 ? lenergy : (henergy - lenergy) * (double)k / ((double)layers - 1.0) + lenergy;
    const long W = spectral || nomap ? n : oversamp * n;

    if (print_only) {
        std::cout.precision(17);
        std::cout << "{\"tbfile\": \"" << json_escape(tbfile) << "\", \"qpifile\": \"" << json_escape(qpifile)
                  << "\", \"name\": \"" << json_escape(tbname) << "\", \"bands\": " << bands
                  << ", \"maxbands\": " << maxbands << ", \"hoppings\": " << nhops << ", \"lattice\": " << n
                  << ", \"kpoints\": " << kpoints << ", \"oversamp\": " << oversamp << ", \"window\": " << window
                  << ", \"layers\": " << layers << ", \"energies\": [" << lenergy << ", " << henergy
                  << "], \"eta\": " << eta << ", \"fermi\": " << efermi << ", \"spin\": " << (spin ? "true" : "false")
                  << ", \"scmodel\": " << (scmodel ? "true" : "false") << ", \"precision\": \"" << precision
                  << "\", \"output\": \""
                  << (josephson ? "josephson" : spectral == 1 ? "spf" : spectral == 2 ? "uspf" : nomap ? "nomode" : "wannier")
                  << "\", \"threshold\": ";
        if (std::isnan(threshold)) std::cout << "null";
        else std::cout << threshold;
        std::cout << ", \"green\": \""
                  << (green == CQPI_GREEN_SURFACE ? "surface" : green == CQPI_GREEN_BULK ? "bulk" : "normal")
                  << "\", \"epserr\": " << epserr << ", \"stbfile\": \"" << json_escape(stbfile)
                  << "\", \"surface_hoppings\": " << nshops;
        std::cout << ", \"deltat\": " << deltat << ", \"etat\": " << etat << ", \"spectral\": " << spectral
                  << ", \"dosfile\": \"" << json_escape(dosfile) << "\", \"doslayers\": " << doslayers
                  << ", \"dosenergies\": [" << doslenergy << ", " << doshenergy << "], \"pos\": [";
        for (size_t i = 0; i < maxbands; i++)
            std::cout << (i ? ", " : "") << "[" << posxy[2 * i] << ", " << posxy[2 * i + 1] << "]";
        std::cout << "], \"vscat\": [";
        for (size_t i = 0; i < vscat.size(); i++)
            std::cout << (i ? ", " : "") << "[" << vscat[i].real() << ", " << vscat[i].imag() << "]";
        double tsum = 0;
        for (double v : tiles) tsum += v;
        std::cout << "], \"spinarr\": ";
        if (!have_spinarr) std::cout << "null";
        else {
            std::cout << "[";
            for (size_t i = 0; i < spinarr.size(); i++) std::cout << (i ? ", " : "") << spinarr[i];
            std::cout << "]";
        }
        std::cout << ", \"tile_sum\": " << tsum << "}\n";
        cqpi_free(hops);
        cqpi_free(shops);
        lf.warn_unused(std::cerr);
        return 0;
    }

    *log << "calcqpi_amd: continuum LDOS on MI355X (" << cqpi_version() << ")\n"
         << "         Tight-binding model  : " << tbfile << " (" << tbname << ", " << bands << " bands)\n"
         << "         Output file for QPI  : " << qpifile << "\n"
         << "         Lattice              : " << n << "x" << n << ", kpoints " << kpoints << "x" << kpoints << "\n"
         << "         energy range         : (" << lenergy << "," << henergy << "), " << layers << " layers\n"
         << "         broadening parameter : " << eta << ", window " << window << ", oversamp " << oversamp << "\n"
         << "         contraction precision: " << precision << "\n";
    if (green != CQPI_GREEN_NORMAL)
        *log << "         Surface tight-binding model: " << stbfile << "\n"
             << "         epserr               : " << epserr << "\n"
             << "         Green's function     : " << (green == CQPI_GREEN_SURFACE ? "surface" : "bulk") << "\n";

    auto t0 = std::chrono::steady_clock::now();
    SectionTimes timings;
    cqpi_ctx *ctx = nullptr;
    if (spectral || nomap) {
        // spectral maps (and the DOS) from a spectral-function context
        if (check(cqpi_create_spectral((size_t)kpoints, (size_t)n, bands, maxbands, spin, posxy.data(), maxbands,
                                       spectral == 2 ? sprearr.data() : nullptr, spectral == 2 ? maxbands : 0, &ctx),
                  "cqpi_create_spectral"))
            return 1;
    } else if (!std::isnan(threshold) && !josephson) {
        // threshold list over the maxbands tiles (src/calcqpi.cc:620-634)
        const size_t wins = (size_t)(2 * window + 1);
        vector<size_t> counts(wins * wins * wins * wins);
        size_t total = 0;
        // with spinarr: the spin-label rule over every orbital (src/qpi.cc:213-253)
        auto build = [&](cqpi_flagentry *out, size_t cap) {
            if (have_spinarr)
                return check(cqpi_make_threshold_list_spinarr(tiles.data(), maxbands, (size_t)window,
                                                              (size_t)oversamp, threshold, spinarr.data(),
                                                              spinarr.size(), out, cap, counts.data(), &total),
                             "cqpi_make_threshold_list_spinarr");
            return check(cqpi_make_threshold_list(tiles.data(), maxbands, (size_t)window, (size_t)oversamp,
                                                  threshold, maxbands, spin, out, cap, counts.data(), &total),
                         "cqpi_make_threshold_list");
        };
        if (build(nullptr, 0)) return 1;
        vector<cqpi_flagentry> flags(std::max<size_t>(total, 1));
        if (build(flags.data(), total)) return 1;
        const size_t entries = wins * wins * wins * wins * (size_t)(oversamp * oversamp) * maxbands * maxbands;
        *log << "Threshold list has " << total << "/" << entries << " entries (" << 100.0 * total / entries
             << "%), requiring " << total * sizeof(cqpi_flagentry) / 1024.0 / 1024.0 / 1024.0 << "GB of memory.\n";
        if (check(cqpi_create_list((size_t)W, (size_t)kpoints, (size_t)n, (size_t)window, (size_t)oversamp, bands,
                                   flags.data(), counts.data(), &ctx),
                  "cqpi_create_list"))
            return 1;
    } else if (check(cqpi_create((size_t)W, (size_t)kpoints, (size_t)n, (size_t)window, (size_t)oversamp, bands,
                                 maxbands, spin, tiles.data(), &ctx),
                     "cqpi_create"))
        return 1;
    char info[2048];
    if (!cqpi_printinfo(ctx, info, sizeof info)) *log << info;
    vector<double> V(2 * bands * bands, 0.0);
    for (size_t i = 0; i < bands; i++) {
        V[2 * (i * bands + i)] = vscat[i].real();
        V[2 * (i * bands + i) + 1] = vscat[i].imag();
    }
    int rc = check(cqpi_set_model(ctx, hops, nhops, efermi, scmodel), "cqpi_set_model");
    if (!rc && green != CQPI_GREEN_NORMAL)
        rc = check(cqpi_set_surface(ctx, green, epserr, shops, nshops), "cqpi_set_surface");
    cqpi_free(hops);
    cqpi_free(shops);
    if (!rc && lf.probe("bsfile")) {
        // band-structure histogram: src/calcqpi.cc:308-326
        long bslattice = 201, bslayers = 21, bsoversamp = 8;
        double bslenergy = -0.1, bshenergy = 0.1;
        const string bsfile = lf.str("bsfile");
        if (lf.probe("bslayers")) bslayers = lf.integer("bslayers");
        if (lf.probe("bsenergies")) {
            vector<double> en = lf.vec("bsenergies");
            if (en.size() >= 2) {
                bslenergy = en[0];
                bshenergy = en[1];
            }
        }
        if (lf.probe("bslattice")) bslattice = lf.integer("bslattice");
        if (lf.probe("bsoversamp")) bsoversamp = lf.integer("bsoversamp");
        *log << "Calculating band structure histogram in energy range (" << bslenergy << "," << bshenergy << ") with "
             << bslayers << " layers and oversampling " << bsoversamp << ".\n";
        IdlMap bs;
        bs.sx = bs.sy = (size_t)bslattice;
        bs.sz = (size_t)bslayers;
        bs.data.assign(bs.sx * bs.sy * bs.sz, 0.0);
        rc = check(cqpi_bands_histogram(ctx, (size_t)bslattice, (size_t)bsoversamp, bslenergy, bshenergy,
                                        (size_t)bslayers, bs.data.data()),
                   "cqpi_bands_histogram");
        if (!rc) {
            // idl(pix, pix, layers, 1, 1, le, ue, 0, -1/2, -1/2), named after the model (src/tightbinding.cc:1161,1191)
            const double h[7] = {1.0, 1.0, -0.5, -0.5, bslenergy, bshenergy, 0.0};
            std::copy(h, h + 7, bs.hdr);
            bs.name = tbname;
            const time_t now = time(nullptr);
            bs.date = ctime(&now);
            if (!bs.date.empty() && bs.date.back() == '\n') bs.date.pop_back();
            if (!write_idl(bsfile, bs)) rc = 1;
            else *log << "... done and saved to " << bsfile << ".\n";
        }
        timings.mark("Bandstructure calculation");
    }
    if (!rc && !dosfile.empty()) {
        // energies lb + k (hb - lb)/layers, k < layers (src/calcqpi.cc:348-356)
        const double step = (doshenergy - doslenergy) / (double)doslayers;
        vector<double> dosw(doslayers), dos(doslayers);
        for (long k = 0; k < doslayers; k++) dosw[k] = doslenergy + (double)k * step;
        *log << "Calculating density of states in energy range (" << doslenergy << "," << doshenergy << ") with "
             << doslayers << " layers.\n";
        rc = check(cqpi_dos(ctx, dosw.data(), dosw.size(), eta, dos.data()), "cqpi_dos");
        if (!rc) {
            std::ofstream dosout(dosfile);
            for (long k = 0; k < doslayers; k++) dosout << dosw[k] << " " << dos[k] << "\n";
            *log << "... done and saved to " << dosfile << ".\n";
        }
        timings.mark("Calculation of density of states");
    }
    if (nomap) {
        cqpi_destroy(ctx);
        if (rc) return 1;
        timings.print(*log);
        lf.warn_unused(std::cerr);
        return 0;
    }
    if (spectral) {
        if (!rc) rc = check(cqpi_set_spectral(ctx, spectral), "cqpi_set_spectral");
    } else {
        if (!rc) rc = check(cqpi_set_impurity(ctx, V.data()), "cqpi_set_impurity");
        if (!rc) rc = check(cqpi_set_precision(ctx, precision == "mixed"  ? CQPI_PREC_MIXED
                                                        : precision == "fp32" ? CQPI_PREC_FP32
                                                                              : CQPI_PREC_FP64), "cqpi_set_precision");
        if (!rc && josephson) rc = check(cqpi_set_josephson(ctx, 1, deltat, etat), "cqpi_set_josephson");
    }
    IdlMap out;
    out.data.assign((size_t)W * W * layers, 0.0);
    timings.mark("Initialization of calculation");
    if (!rc) rc = check(cqpi_run(ctx, omegas.data(), omegas.size(), eta, out.data.data()), "cqpi_run");
    cqpi_destroy(ctx);
    if (rc) return 1;
    timings.mark("Calculation of CLDOS");
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    // src/calcqpi.cc:639-652
    std::ostringstream comment;
    if (spectral == 1) comment << "SPF calculation for " << tbname;
    else if (spectral == 2) comment << "USPF calculation for " << tbname << " with kpoints=" << kpoints;
    else
        comment << (josephson ? "Josephson calculation for " : "CQPI calculation for ") << tbname
                << " with window=" << window << " and oversamp=" << oversamp;
    out.name = comment.str();
    time_t now = time(nullptr);
    out.date = ctime(&now);
    if (!out.date.empty() && out.date.back() == '\n') out.date.pop_back();
    out.sx = out.sy = (size_t)W;
    out.sz = (size_t)layers;
    // idl(W,W,L, 2.0,2.0, lb,hb, 0.0, -1.0,-1.0) in file order: scan, offset, bias, setpoint
    double hdr[7] = {2.0, 2.0, -1.0, -1.0, lenergy, henergy, 0.0};
    std::copy(hdr, hdr + 7, out.hdr);
    *log << "Calculation finished in " << secs << " s (" << (double)W * W * layers / secs
         << " pixel-energy points/s including set-up and copies).\nSaving to " << qpifile << "\n";
    if (!write_idl(qpifile, out)) {
        std::cerr << "Error writing " << qpifile << "\n";
        return 1;
    }
    *log << "Done.\n";
    timings.mark("Saving data");
    timings.print(*log);
    lf.warn_unused(std::cerr);
    return 0;
}
