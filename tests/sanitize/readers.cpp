// The host readers (control files, SpecMesh meshes, .hsol files) driven over good and malformed
// inputs, for a build with AddressSanitizer and UndefinedBehaviorSanitizer.  argv[1]: a folder
// with case_<name>/{case.control, box.mesh, orders.txt}, bad/*.control, bad/trunc_*.mesh,
// bad/trunc_*_<hsol>.hsol and <hsol>.N (element orders); argv[2]: the .hsol fixture folder.
// Every result goes to stdout, one line each, for the calling test to check.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>
#include "../../include/hx3d.h"

static const char *ZONES[] = {"left", "top", "back", "bottom", "front", "right", "nowhere", ""};
static const char *KEYS[] = {"mesh file name", "MACH NUMBER", "reynolds number", "cfl", "dt", "polynomial order",
                             "", "averaging", "x_not_a_key", nullptr};

static void control(const std::string &path, const std::string &root)
{
    hx3d_control *c = nullptr;
    int rc = hx3d_control_read(path.c_str(), &c);
    std::printf("control %s rc %d %s\n", path.c_str(), rc, c ? (hx3d_control_error(c) ? hx3d_control_error(c) : "") : "(null)");
    if (!c) return;
    for (int k = 0; KEYS[k]; ++k) {
        const char *v = hx3d_control_value(c, KEYS[k]);
        std::printf("  [%s] = %s\n", KEYS[k], v ? v : "(none)");
    }
    for (const char *z : ZONES) {
        char t[32], cp[32];
        for (int len : {32, 4, 1}) {
            int r = hx3d_control_boundary(c, z, t, cp, len);
            std::printf("  zone %s len %d rc %d [%s] [%s]\n", z, len, r, r == 0 ? t : "", r == 0 ? cp : "");
        }
    }
    hx3d_mesh *m = nullptr;
    hx3d_case_desc d;
    std::memset(&d, 0, sizeof d);
    hx3d_control_run run;
    rc = hx3d_control_case(c, root.empty() ? nullptr : root.c_str(), &m, &d, &run);
    std::printf("  case rc %d %s K %d nodes %ld\n", rc, hx3d_control_error(c) ? hx3d_control_error(c) : "",
                rc == 0 ? d.K : -1, m ? hx3d_mesh_num_nodes(m) : -1L);
    if (m) hx3d_mesh_destroy(m);
    hx3d_control_destroy(c);
}

static void meshfile(const std::string &path)
{
    for (const char *per : {"left=right;front=back;bottom=top", "", "left=right", "a=b;;=;x", (const char *)nullptr}) {
        hx3d_mesh *m = nullptr;
        int rc = hx3d_mesh_read_specmesh(path.c_str(), 3, nullptr, per, &m);
        std::printf("mesh %s [%s] rc %d %s\n", path.c_str(), per ? per : "null", rc, m && hx3d_mesh_error(m) ? hx3d_mesh_error(m) : "");
        if (m && rc == 0) {
            hx3d_case_desc d;
            std::memset(&d, 0, sizeof d);
            hx3d_mesh_fill_desc(m, &d);
            double s = 0;
            const double *x = hx3d_mesh_coords(m);
            for (long g = 0; g < 3 * hx3d_mesh_num_nodes(m); ++g) s += x[g];
            std::printf("  K %d nodes %ld fnodes %ld sum %.6f\n", d.K, hx3d_mesh_num_nodes(m), hx3d_mesh_num_face_nodes(m), s);
        }
        if (m) hx3d_mesh_destroy(m);
    }
    // orders arrays of the wrong range
    std::vector<int> ord(3 * 8, 0);
    hx3d_mesh *m = nullptr;
    int rc = hx3d_mesh_read_specmesh(path.c_str(), 3, ord.data(), "left=right;front=back;bottom=top", &m);
    std::printf("mesh zero orders rc %d %s\n", rc, m && hx3d_mesh_error(m) ? hx3d_mesh_error(m) : "");
    if (m) hx3d_mesh_destroy(m);
}

static void hsol(const std::string &path, const std::vector<int> &N, const std::string &out)
{
    int type = 0, nt = 0, K = 0, iter = 0;
    double time = 0, refs[6];
    int rc = hx3d_hsol_read_header(path.c_str(), &type, &nt, &K, &iter, &time, refs);
    std::printf("hsol %s header rc %d type %d K %d iter %d %s\n", path.c_str(), rc, type, K, iter, rc ? hx3d_hsol_error() : "");
    if (rc) return;
    long nodes = 0;
    for (size_t e = 0; e + 2 < N.size(); e += 3) nodes += (long)(N[e] + 1) * (N[e + 1] + 1) * (N[e + 2] + 1);
    std::vector<double> Q(5 * nodes), G(15 * nodes);
    rc = hx3d_hsol_read(path.c_str(), (int)N.size() / 3, N.data(), Q.data(), type == 3 ? G.data() : nullptr);
    std::printf("  read rc %d %s\n", rc, rc ? hx3d_hsol_error() : "");
    std::vector<int> wrong(N);
    if (!wrong.empty()) wrong[0] += 1;
    rc = hx3d_hsol_read(path.c_str(), (int)N.size() / 3, wrong.data(), Q.data(), nullptr);
    std::printf("  read wrong orders rc %d %s\n", rc, rc ? hx3d_hsol_error() : "");
    if (type && K == (int)N.size() / 3) {
        rc = hx3d_hsol_write(out.c_str(), nullptr, K, N.data(), iter, time, refs, Q.data(), type == 3 ? G.data() : nullptr);
        std::vector<double> Q2(Q.size());
        int rc2 = hx3d_hsol_read(out.c_str(), K, N.data(), Q2.data(), nullptr);
        std::printf("  roundtrip rc %d %d equal %d\n", rc, rc2, (int)(Q == Q2));
    }
}

static std::vector<int> readN(const std::string &p)
{
    std::ifstream f(p);
    std::vector<int> v;
    int x;
    while (f >> x) v.push_back(x);
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    const std::string W = std::string(argv[1]) + "/";
    for (const char *n : {"order_file_euler_pirozzoli", "orders_ijk_smagorinsky", "syntax_split_kg_lxf"}) {
        control(W + "case_" + n + "/case.control", W + "case_" + n);
        control(W + "case_" + n + "/case.control", "");
        meshfile(W + "case_" + n + "/box.mesh");
    }
    for (const char *b : {"empty", "noeq", "longkey", "quote", "define", "eqonly", "nomesh", "badnum", "missing"})
        control(W + "bad/" + b + ".control", W + "case_syntax_split_kg_lxf");
    for (const char *f : {"0", "0.01", "0.1", "0.5", "0.9", "0.999"}) meshfile(W + "bad/trunc_" + f + ".mesh");
    meshfile(W + "bad/none.mesh");
    const std::string H = std::string(argv[2]) + "/";
    for (const char *c : {"box2_mixed_p2to4", "box2_p3_curved"}) {
        std::vector<int> N = readN(W + c + ".N");
        for (const char *s : {"_step0", "_final"}) {
            hsol(H + c + s + ".hsol", N, W + "roundtrip.hsol");
            for (const char *f : {"0", "0.05", "0.5", "0.99"}) hsol(W + "bad/trunc_" + f + "_" + c + s + ".hsol", N, W + "roundtrip.hsol");
        }
    }
    std::printf("done\n");
    return 0;
}
