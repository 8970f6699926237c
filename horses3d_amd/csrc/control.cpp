// HORSES3D .control files for the hx3d engine: the reference's reader and the keys that set up
// the compressible Navier-Stokes time derivative and its explicit RK3 loop.
//
//   main dictionary        ReadControlFile (ReadInputFile.f90:20-125) with PreprocessInputLine,
//                          GetKeyword, GetValueAsString (FileReading.f90:10-82); keys live in an
//                          FTValueDictionary (FTDictionaryClass.f90:371-535: the first entry
//                          added under a key is the one found; keys hold 64 characters)
//   defaults, checks       CheckInputIntegrity (NavierStokesSolver/main.f90:239-318), the
//                          physics keys of ConstructPhysicsStorage_NS (PhysicsStorage_NS.f90:
//                          186-240, 525-540)
//   boundary blocks        GetZoneType (BoundaryConditions.f90:170-260), ConstructPeriodicBC
//                          (PeriodicBC.f90:73-151), CheckIfBoundaryNameIsContained
//                          (GenericBoundaryConditionClass.f90:390-458)
//   selectors              SetRiemannSolver (RiemannSolvers_NS.f90:120-287), the inviscid and
//                          viscous discretizations (SpatialDiscretization.f90:80-175), the LES
//                          model (LESModels.f90:85-125, 240-250, 344-354), the node type
//                          (NodalStorageClass.f90:83-103)
//   orders                 GetMeshPolynomialOrders (pAdaptationClass.f90:198-240), ReadOrderFile
//                          (ReadInputFile.f90:132-153)
//   time integrator        constructTimeIntegrator (TimeIntegrator.f90:178-214)
//
// Host code only.  A number that does not read is an error here (the reference's dictionary
// gives a NaN and the run carries it on).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hx3d.h"
#include "frontend.hpp"

namespace {

constexpr size_t LINE_LENGTH = 132;   // character(len=LINE_LENGTH) lines, '(A132)' records
constexpr size_t KEY_LENGTH = 64;     // FTDICT_KWD_STRING_LENGTH

// Fortran blanks are spaces: ADJUSTL and TRIM move and drop spaces only
std::string rtrim(const std::string &s)
{
    size_t e = s.find_last_not_of(' ');
    return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}
std::string adjustl(const std::string &s)
{
    size_t b = s.find_first_not_of(' ');
    return b == std::string::npos ? std::string() : s.substr(b);
}
std::string lower(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// PreprocessInputLine: everything from the first '!' on is a comment; a line starting with '/'
// is blank
void preprocess(std::string &line)
{
    const size_t p = line.find('!');
    if (p != std::string::npos) line.resize(p);
    if (!line.empty() && line[0] == '/') line.clear();
}

// GetKeyword: the text before the first '=' of the left-adjusted line (empty without '=')
std::string get_keyword(const std::string &line)
{
    const std::string a = adjustl(line);
    const size_t eq = a.find('=');
    return eq == std::string::npos ? std::string() : a.substr(0, eq);
}

// GetValueAsString: the text between the first and the last '"', else everything after the
// first '='
std::string get_value(const std::string &line)
{
    const size_t q1 = line.find('"');
    if (q1 != std::string::npos) {
        const size_t q2 = line.rfind('"');
        return q2 > q1 ? line.substr(q1 + 1, q2 - q1 - 1) : std::string();
    }
    const size_t eq = line.find('=');
    const std::string t = rtrim(line);
    const size_t from = eq == std::string::npos ? 0 : eq + 1;
    return from < t.size() ? t.substr(from) : std::string();
}

struct Dict {
    std::vector<std::pair<std::string, std::string>> kv;   // insertion order
    void add(const std::string &key, const std::string &value)
    {
        kv.emplace_back(rtrim(key.substr(0, std::min(key.size(), KEY_LENGTH))), rtrim(value));
    }
    const std::string *find(const std::string &key) const
    {
        const std::string k = rtrim(key);
        for (const auto &p : kv)
            if (p.first == k) return &p.second;
        return nullptr;
    }
};

// CheckIfBoundaryNameIsContained: a header "#define boundary a__b__c" (from column 1) lists the
// zones it applies to, separated by "__"
bool boundary_name_contained(const std::string &line, const std::string &bname)
{
    static const std::string head = "#define boundary";
    if (line.compare(0, head.size(), head) != 0) return false;
    const std::string str = rtrim(adjustl(rtrim(line).substr(std::min(line.size(), head.size()))));
    if (str.empty()) return false;
    size_t pos1 = 0;
    while (true) {
        const size_t p = str.find("__", pos1);
        const std::string cur = p == std::string::npos ? str.substr(pos1) : str.substr(pos1, p - pos1);
        if (rtrim(cur) == rtrim(bname)) return true;
        if (p == std::string::npos) return false;
        pos1 = p + 2;
    }
}

// list-directed input of one value: the first item of the record, ended by a blank, ',' or '/'
std::string first_item(const std::string &s)
{
    const size_t b = s.find_first_not_of(" \t");
    if (b == std::string::npos) return std::string();
    const size_t e = s.find_first_of(" \t,/", b);
    return s.substr(b, e == std::string::npos ? std::string::npos : e - b);
}

bool read_real(const std::string &s, double &v)
{
    std::string t = first_item(s);
    if (t.empty()) return false;
    for (auto &c : t)
        if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'e';
    char *end = nullptr;
    errno = 0;
    v = std::strtod(t.c_str(), &end);
    return end && *end == '\0' && errno == 0 && end != t.c_str();
}

bool read_int(const std::string &s, long &v)
{
    const std::string t = first_item(s);
    if (t.empty()) return false;
    char *end = nullptr;
    errno = 0;
    v = std::strtol(t.c_str(), &end, 10);
    return end && *end == '\0' && errno == 0 && end != t.c_str();
}

// FTValue logicalValue of a string: only .true./.false. (either case) read as logicals
bool read_logical(const std::string &s)
{
    const std::string t = rtrim(s);
    return t == ".true." || t == ".TRUE.";
}

bool is_absolute(const std::string &p) { return !p.empty() && p[0] == '/'; }

}  // namespace

struct hx3d_control {
    std::string err, path;
    std::vector<std::string> lines;   // the file's records, cut to LINE_LENGTH
    Dict dict;                        // the main dictionary, with CheckInputIntegrity's defaults
};

namespace {

int fail(hx3d_control *c, int code, const std::string &m)
{
    c->err = m;
    return code;
}

// the keys one boundary zone's blocks set (GetZoneType / ConstructPeriodicBC): lines lowered,
// a block is entered at a "#define boundary" header naming the zone, and the first "#end"
// inside one ends the search
Dict zone_dict(const hx3d_control *c, const std::string &zone)
{
    Dict d;
    bool inside = false;
    const std::string bname = lower(zone);
    for (std::string line : c->lines) {
        preprocess(line);
        line = lower(line);
        if (rtrim(line).find("#define boundary") != std::string::npos)
            inside = boundary_name_contained(rtrim(line), bname);
        if (!inside) continue;
        if (rtrim(line) == "#end") break;
        d.add(lower(adjustl(get_keyword(line))), adjustl(get_value(line)));
    }
    return d;
}

int zone_boundary(const hx3d_control *c, const std::string &zone, std::string &type, std::string &coupled,
                  std::string &err)
{
    static const char *implemented[] = {"inflow",   "outflow",      "noslipwall",     "freeslipwall",
                                        "periodic", "user-defined", "manufacturedsol", "msoutflowspecifyp"};
    const Dict d = zone_dict(c, zone);
    const std::string *t = d.find("type");
    if (!t) {
        err = "Missing boundary condition type for boundary " + zone;
        return -HX3D_EINVAL;
    }
    type = lower(rtrim(*t));
    bool known = false;
    for (const char *n : implemented) known = known || type == n;
    if (!known) {
        err = "Boundary type " + type + " not recognized";
        return -HX3D_EINVAL;
    }
    if (type != "periodic") return 0;
    const std::string *cb = d.find("coupled boundary");
    if (!cb) {
        err = "Select a \"coupled boundary\" for boundary " + zone;
        return -HX3D_EINVAL;
    }
    coupled = rtrim(*cb);
    return 1;
}

int need_real(hx3d_control *c, const char *key, double &v)
{
    const std::string *s = c->dict.find(key);
    if (!s) return fail(c, HX3D_EINVAL, std::string("Input file is missing entry for keyword: ") + key);
    if (!read_real(*s, v)) return fail(c, HX3D_EINVAL, std::string("\"") + key + "\" = " + *s + " is not a number");
    return HX3D_OK;
}

int need_int(hx3d_control *c, const char *key, long &v)
{
    const std::string *s = c->dict.find(key);
    if (!s) return fail(c, HX3D_EINVAL, std::string("Input file is missing entry for keyword: ") + key);
    if (!read_int(*s, v)) return fail(c, HX3D_EINVAL, std::string("\"") + key + "\" = " + *s + " is not an integer");
    return HX3D_OK;
}

int read_control(hx3d_control *c, const char *path)
{
    std::ifstream in(path);
    if (!in) return fail(c, HX3D_EINVAL, std::string("Wrong control file \"") + path + "\"");
    c->path = path;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() > LINE_LENGTH) line.resize(LINE_LENGTH);
        c->lines.push_back(line);
    }
    // ReadControlFile: "#define ... #end" blocks belong to other readers
    bool hashtag = false;
    for (std::string l : c->lines) {
        preprocess(l);
        if (rtrim(l).empty()) continue;
        if (l.find("#define") != std::string::npos) {
            hashtag = true;
        } else if (l.find("#end") != std::string::npos && hashtag) {
            hashtag = false;
            continue;
        }
        if (hashtag) continue;
        c->dict.add(lower(adjustl(get_keyword(l))), adjustl(get_value(l)));
    }
    // CheckInputIntegrity's defaults
    const char *defaults[][2] = {{"save gradients with solution", ".false."},
                                 {"save sensor with solution", ".false."},
                                 {"discretization nodes", "Gauss"},
                                 {"inviscid discretization", "Standard"},
                                 {"viscous discretization", "BR1"},
                                 {"split form", "Ducros"}};
    for (auto &d : defaults)
        if (!c->dict.find(d[0])) c->dict.add(d[0], d[1]);
    // split-form discretizations run on Gauss-Lobatto nodes
    if (lower(*c->dict.find("inviscid discretization")) == "split-form" &&
        lower(*c->dict.find("discretization nodes")) == "gauss") {
        for (auto &p : c->dict.kv)
            if (p.first == "discretization nodes") {
                p.second = "Gauss-Lobatto";
                break;
            }
    }
    return HX3D_OK;
}

int fill_case(hx3d_control *c, const char *root, hx3d_mesh **mesh, hx3d_case_desc *d, hx3d_control_run *run)
{
    // required keys (mainKeywords, Setup.f90:21-32; physics_NSKeywords, PhysicsStorage_NS.f90:21)
    static const char *required[] = {"mesh file name",        "inviscid discretization",     "split form",
                                     "discretization nodes",  "save gradients with solution",
                                     "save sensor with solution", "restart",                   "restart file name",
                                     "solution file name",    "number of time steps",        "output interval",
                                     "convergence tolerance", "mach number",                 "flow equations"};
    for (const char *k : required)
        if (!c->dict.find(k)) return fail(c, HX3D_EINVAL, std::string("Input file is missing entry for keyword: ") + k);
    auto str = [c](const char *k) { return c->dict.find(k) ? rtrim(*c->dict.find(k)) : std::string(); };
    auto path_of = [root](const std::string &p) {
        return (root && *root && !is_absolute(p)) ? std::string(root) + "/" + p : p;
    };

    hx3d_case_desc desc;
    std::memset(&desc, 0, sizeof desc);
    int rc;
    // physics: ConstructPhysicsStorage_NS
    if ((rc = need_real(c, "mach number", desc.mach))) return rc;
    desc.prandtl = 0.72;
    if (c->dict.find("prandtl number") && (rc = need_real(c, "prandtl number", desc.prandtl))) return rc;
    desc.euler = lower(str("flow equations")) == "euler";
    if (!desc.euler && (rc = need_real(c, "reynolds number", desc.reynolds))) return rc;

    // discretization nodes: the names are matched as written (InitializeNodalStorage_controlVars,
    // NodalStorageClass.f90:80-105)
    const std::string nodes = str("discretization nodes");
    if (nodes == "Gauss") desc.nodes = HX3D_NODES_GAUSS;
    else if (nodes == "Gauss-Lobatto") desc.nodes = HX3D_NODES_GAUSSLOBATTO;
    else return fail(c, HX3D_EINVAL, "Unknown discretization nodes \"" + nodes + "\"");
    const bool gauss = desc.nodes == HX3D_NODES_GAUSS;

    const std::string inv = lower(str("inviscid discretization"));
    if (inv == "standard") desc.inviscid = HX3D_INVISCID_STANDARD;
    else if (inv == "split-form") desc.inviscid = HX3D_INVISCID_SPLIT;
    else return fail(c, HX3D_EINVAL, "Requested inviscid discretization \"" + inv + "\" is not implemented");

    if (!desc.euler) {
        const std::string vis = lower(str("viscous discretization"));
        if (vis == "br2" || vis == "ip")
            return fail(c, HX3D_EUNSUPPORTED, "viscous discretization \"" + vis + "\": the engine runs BR1");
        if (vis != "br1") return fail(c, HX3D_EINVAL, "Viscous discretization \"" + vis + "\" not recognized");
    }

    // SetRiemannSolver
    desc.riemann = HX3D_RIEMANN_ROE;
    if (c->dict.find("riemann solver")) {
        const std::string r = lower(str("riemann solver"));
        static const char *other[] = {"central",     "rusanov",          "standard roe", "roe-pike",
                                      "low dissipation roe", "matrix dissipation", "viscous ns", "u-diss"};
        if (r == "roe") desc.riemann = HX3D_RIEMANN_ROE;
        else if (r == "lax-friedrichs") desc.riemann = HX3D_RIEMANN_LXF;
        else {
            for (const char *o : other)
                if (r == o) return fail(c, HX3D_EUNSUPPORTED, "riemann solver \"" + r + "\": the engine runs Roe and Lax-Friedrichs");
            return fail(c, HX3D_EINVAL, "Riemann Solver not recognized: \"" + r + "\"");
        }
    }
    desc.lambda_stab = 1.0;
    if (c->dict.find("lambda stabilization") && (rc = need_real(c, "lambda stabilization", desc.lambda_stab)))
        return rc;
    desc.averaging = HX3D_AVG_STANDARD;
    if (c->dict.find("averaging")) {
        const std::string a = lower(str("averaging"));
        if (a == "standard") desc.averaging = HX3D_AVG_STANDARD;
        else if (a == "ducros") desc.averaging = HX3D_AVG_DUCROS;
        else if (a == "kennedy-gruber") desc.averaging = HX3D_AVG_KENNEDYGRUBER;
        else if (a == "pirozzoli") desc.averaging = HX3D_AVG_PIROZZOLI;
        else if (a == "morinishi") desc.averaging = HX3D_AVG_MORINISHI;
        else if (a == "entropy conserving" || a == "chandrasekar")
            return fail(c, HX3D_EUNSUPPORTED, "averaging \"" + a + "\" is not supported by the engine");
        else return fail(c, HX3D_EINVAL, "Averaging not recognized: \"" + a + "\"");
    }

    // LES model (read for Navier-Stokes flows only)
    desc.les = HX3D_LES_NONE;
    if (!desc.euler && c->dict.find("les model")) {
        const std::string m = lower(str("les model"));
        if (m == "smagorinsky") desc.les = HX3D_LES_SMAGORINSKY, desc.les_c = 0.2;
        else if (m == "wale") desc.les = HX3D_LES_WALE, desc.les_c = 0.325;
        else if (m == "vreman") return fail(c, HX3D_EUNSUPPORTED, "LES Model vreman is not supported by the engine");
        else if (m != "none") return fail(c, HX3D_EINVAL, "LES Model " + m + " is not implemented");
        if (desc.les != HX3D_LES_NONE && c->dict.find("les model intensity") &&
            (rc = need_real(c, "les model intensity", desc.les_c)))
            return rc;
    }

    // time integrator
    hx3d_control_run r;
    std::memset(&r, 0, sizeof r);
    if (c->dict.find("cfl")) {
        if ((rc = need_real(c, "cfl", r.cfl))) return rc;
        if (!desc.euler) {
            if (!c->dict.find("dcfl"))
                return fail(c, HX3D_EINVAL, "\"cfl\" and \"dcfl\", or \"dt\" keyword must be specified for the time integrator");
            if ((rc = need_real(c, "dcfl", r.dcfl))) return rc;
        }
    } else if (c->dict.find("dt")) {
        if ((rc = need_real(c, "dt", r.dt))) return rc;
    } else {
        return fail(c, HX3D_EINVAL, "\"cfl\" or \"dt\" keywords must be specified for the time integrator");
    }
    long v;
    if ((rc = need_int(c, "number of time steps", v))) return rc;
    r.steps = (int)v;
    if ((rc = need_int(c, "output interval", v))) return rc;
    r.output_interval = (int)v;
    if ((rc = need_real(c, "convergence tolerance", r.tolerance))) return rc;
    r.restart = read_logical(str("restart"));

    // polynomial orders: an order file, one order, or one order per direction
    std::vector<int> orders;
    int Nuni = 0;
    if (c->dict.find("polynomial order file")) {
        std::ifstream of(path_of(str("polynomial order file")));
        if (!of) return fail(c, HX3D_EINVAL, "cannot open polynomial order file " + str("polynomial order file"));
        long ne = 0;
        if (!(of >> ne) || ne <= 0) return fail(c, HX3D_EINVAL, "bad polynomial order file");
        orders.resize(3 * ne);
        std::string rest;
        std::getline(of, rest);   // list-directed: one READ per record
        for (long e = 0; e < ne; ++e) {
            std::string rec, all;
            int got = 0;
            long o[3];
            while (got < 3 && std::getline(of, rec)) {
                for (auto &ch : rec)
                    if (ch == ',') ch = ' ';
                char *p = &rec[0];
                while (got < 3) {
                    char *end;
                    const long x = std::strtol(p, &end, 10);
                    if (end == p) break;
                    o[got++] = x;
                    p = end;
                }
            }
            if (got < 3) return fail(c, HX3D_EINVAL, "polynomial order file is truncated");
            for (int k = 0; k < 3; ++k) orders[3 * e + k] = (int)o[k];
        }
    } else if (c->dict.find("polynomial order")) {
        if ((rc = need_int(c, "polynomial order", v))) return rc;
        Nuni = (int)v;
    } else if (c->dict.find("polynomial order i") && c->dict.find("polynomial order j") &&
               c->dict.find("polynomial order k")) {
        long o[3];
        if ((rc = need_int(c, "polynomial order i", o[0])) || (rc = need_int(c, "polynomial order j", o[1])) ||
            (rc = need_int(c, "polynomial order k", o[2])))
            return rc;
        Nuni = -1;   // one (Nx, Ny, Nz) row for every element
        orders.assign(3, 0);
        for (int k = 0; k < 3; ++k) orders[k] = (int)o[k];
    } else {
        return fail(c, HX3D_EINVAL, "The polynomial order(s) must be specified");
    }

    // the mesh, with the boundary conditions of the .control file's boundary blocks
    auto resolve = [c](const std::string &zone, std::string &coupled, std::string &err) {
        std::string type;
        return zone_boundary(c, zone, type, coupled, err);
    };
    const std::string mfile = path_of(str("mesh file name"));
    hx3d_mesh *m = nullptr;
    if (Nuni == -1)
        rc = hx3d_front::read_specmesh(&m, mfile.c_str(), 0, orders.data(), hx3d_front::ORDERS_ONE_ROW, resolve, gauss);
    else if (!orders.empty())
        rc = hx3d_front::read_specmesh(&m, mfile.c_str(), 0, orders.data(), (long)orders.size() / 3, resolve, gauss);
    else
        rc = hx3d_front::read_specmesh(&m, mfile.c_str(), Nuni, nullptr, hx3d_front::ORDERS_UNCHECKED, resolve, gauss);
    *mesh = m;
    if (rc) return fail(c, rc, m ? hx3d_mesh_error(m) : "out of host memory");
    hx3d_mesh_fill_desc(m, &desc);
    *d = desc;
    if (run) *run = r;
    return HX3D_OK;
}

}  // namespace

extern "C" {

int hx3d_control_read(const char *path, hx3d_control **out)
{
    if (!out) return HX3D_EINVAL;
    *out = nullptr;
    hx3d_control *c = new (std::nothrow) hx3d_control;
    if (!c) return HX3D_ENOMEM;
    *out = c;
    if (!path) return fail(c, HX3D_EINVAL, "no control file name");
    try {
        return read_control(c, path);
    } catch (const std::bad_alloc &) {
        return fail(c, HX3D_ENOMEM, "out of host memory reading the control file");
    }
}

const char *hx3d_control_error(const hx3d_control *c) { return c ? c->err.c_str() : "no control file"; }

const char *hx3d_control_value(const hx3d_control *c, const char *key)
{
    if (!c || !key) return nullptr;
    // the dictionary does not change once read: its strings stay where they are
    const std::string *v = c->dict.find(lower(key));
    return v ? v->c_str() : nullptr;
}

int hx3d_control_boundary(hx3d_control *c, const char *zone, char *type, char *coupled, int len)
{
    if (!c || !zone) return HX3D_EINVAL;
    std::string t, cb, err;
    int r;
    try {
        r = zone_boundary(c, zone, t, cb, err);
    } catch (const std::bad_alloc &) {
        return fail(c, HX3D_ENOMEM, "out of host memory");
    }
    if (r < 0) return fail(c, -r, err);
    auto put = [len](char *dst, const std::string &s) {
        if (!dst || len <= 0) return;
        const size_t n = std::min(s.size(), (size_t)len - 1);
        std::memcpy(dst, s.data(), n);
        dst[n] = '\0';
    };
    put(type, t);
    put(coupled, r == 1 ? cb : std::string());
    return HX3D_OK;
}

int hx3d_control_case(hx3d_control *c, const char *root, hx3d_mesh **mesh, hx3d_case_desc *desc,
                      hx3d_control_run *run)
{
    if (!c || !mesh || !desc) return c ? fail(c, HX3D_EINVAL, "hx3d_control_case: missing argument") : HX3D_EINVAL;
    *mesh = nullptr;
    try {
        return fill_case(c, root, mesh, desc, run);
    } catch (const std::bad_alloc &) {
        return fail(c, HX3D_ENOMEM, "out of host memory");
    }
}

int hx3d_control_destroy(hx3d_control *c)
{
    delete c;
    return HX3D_OK;
}

}  // extern "C"
