// HORSES3D solution files (.hsol) of the hx3d engine: the unformatted stream files of
// SolutionFile.f90 that HexMesh_SaveSolution (HexMesh.f90:3376-3556) writes and
// HexMesh_LoadSolution (HexMesh.f90:4582-4912) reads for restarts.
//
//   byte   0  file name (128 characters, the name's own length, zero padded)    POS_FILENAME
//        128  int32 file type (2: solution, 3: solution and gradients)          POS_FILETYPE
//        132  int32 node type (1: Gauss, 2: Gauss-Lobatto)                      POS_NODETYPE
//        136  int32 number of elements                                          POS_NOOFELEMENTS
//        140  int32 iteration                                                   POS_ITER
//        144  float64 time                                                      POS_TIME
//        152  6 x float64 reference values gamma, R, V, rho, T, Mach            POS_REFS
//        200  int32 88                                                          BEGINNING_DATA
//        204  per element: int32 4, int32 5, Nx+1, Ny+1, Nz+1, then Q(1:5, i, j, k) as float64
//             (variable fastest) and, for type 3, U_x, U_y, U_z in the same layout
//        end  int32 99                                                          END_OF_FILE
// (SolutionFile.f90:61-134 positions, :553-570 Write4DArray, :197-213 SealSolutionFile.)
// Little-endian, as the reference writes it on x86-64 and as this host reads it.  Host only.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hx3d.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &m)
{
    g_err = m;
    return code;
}

constexpr int SOLUTION_FILE = 2, SOLUTION_AND_GRADIENTS_FILE = 3, GAUSS = 1, GAUSSLOBATTO = 2;
constexpr int BEGINNING_DATA = 88, END_OF_FILE = 99, NAME_LEN = 128, POS_INIT_DATA = 204;

struct File {
    std::FILE *f = nullptr;
    explicit File(std::FILE *x) : f(x) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

template <class T>
bool put(std::FILE *f, const T &v)
{
    return std::fwrite(&v, sizeof(T), 1, f) == 1;
}
template <class T>
bool get(std::FILE *f, T &v)
{
    return std::fread(&v, sizeof(T), 1, f) == 1;
}

}  // namespace

extern "C" {

const char *hx3d_hsol_error(void) { return g_err.c_str(); }

int hx3d_hsol_reference_values(double mach, double refs[6])
{
    // ThermodynamicsAir and the default reference state of ConstructPhysicsStorage_NS
    // (PhysicsStorage_NS.f90:115-132, :292-306): T = 520 R, p = 101325 Pa
    if (!refs) return fail(HX3D_EINVAL, "no output array");
    const double gamma = 1.4, R = 287.15;
    const double T = 520.0 * 5.0 / 9.0;
    const double rho = 101325.0 / (R * T);
    const double V = mach * std::sqrt(gamma * R * T);
    refs[0] = gamma;
    refs[1] = R;
    refs[2] = V;
    refs[3] = rho;
    refs[4] = T;
    refs[5] = mach;
    return HX3D_OK;
}

int hx3d_hsol_write(const char *path, const char *header_name, int K, const int *N, int iter, double time,
                    const double refs[6], const double *Q, const double *grad)
{
    return hx3d_hsol_write_nodes(path, header_name, K, N, iter, time, refs, Q, grad, HX3D_NODES_GAUSSLOBATTO);
}

int hx3d_hsol_write_nodes(const char *path, const char *header_name, int K, const int *N, int iter, double time,
                          const double refs[6], const double *Q, const double *grad, int node_kind)
{
    if (!path || !N || !Q || !refs || K <= 0) return fail(HX3D_EINVAL, "hx3d_hsol_write: missing argument");
    if (node_kind != HX3D_NODES_GAUSSLOBATTO && node_kind != HX3D_NODES_GAUSS) return fail(HX3D_EINVAL, "unknown node type");
    // the file's node type field: the reference's GAUSS = 1, GAUSSLOBATTO = 2 (NodalStorageClass.f90:16-17)
    const int32_t node_type = node_kind == HX3D_NODES_GAUSS ? GAUSS : GAUSSLOBATTO;
    long nodes = 0;
    for (int e = 0; e < K; ++e) {
        if (N[3 * e] < 0 || N[3 * e + 1] < 0 || N[3 * e + 2] < 0) return fail(HX3D_EINVAL, "negative order");
        nodes += (long)(N[3 * e] + 1) * (N[3 * e + 1] + 1) * (N[3 * e + 2] + 1);
    }
    File f(std::fopen(path, "wb"));
    if (!f.f) return fail(HX3D_EINVAL, std::string("cannot create ") + path);
    char name[NAME_LEN];
    std::memset(name, 0, sizeof name);
    const std::string nm = header_name ? header_name : path;
    std::memcpy(name, nm.data(), std::min<size_t>(nm.size(), NAME_LEN));
    bool ok = std::fwrite(name, 1, NAME_LEN, f.f) == (size_t)NAME_LEN;
    const int32_t type = grad ? SOLUTION_AND_GRADIENTS_FILE : SOLUTION_FILE;
    ok = ok && put(f.f, type) && put(f.f, node_type) && put(f.f, (int32_t)K) && put(f.f, (int32_t)iter);
    ok = ok && put(f.f, time);
    for (int r = 0; r < 6; ++r) ok = ok && put(f.f, refs[r]);
    ok = ok && put(f.f, (int32_t)BEGINNING_DATA);
    long off = 0;
    for (int e = 0; e < K && ok; ++e) {
        const long ne = (long)(N[3 * e] + 1) * (N[3 * e + 1] + 1) * (N[3 * e + 2] + 1);
        ok = put(f.f, (int32_t)4) && put(f.f, (int32_t)5) && put(f.f, (int32_t)(N[3 * e] + 1)) &&
             put(f.f, (int32_t)(N[3 * e + 1] + 1)) && put(f.f, (int32_t)(N[3 * e + 2] + 1));
        ok = ok && std::fwrite(Q + 5 * off, sizeof(double), 5 * ne, f.f) == (size_t)(5 * ne);
        if (grad)
            for (int d = 0; d < 3 && ok; ++d)
                ok = std::fwrite(grad + 5 * (d * nodes + off), sizeof(double), 5 * ne, f.f) == (size_t)(5 * ne);
        off += ne;
    }
    ok = ok && put(f.f, (int32_t)END_OF_FILE);
    if (!ok) return fail(HX3D_EINVAL, std::string("write error on ") + path);
    return HX3D_OK;
}

int hx3d_hsol_read_header(const char *path, int *type, int *node_type, int *K, int *iter, double *time, double refs[6])
{
    if (!path) return fail(HX3D_EINVAL, "no file name");
    File f(std::fopen(path, "rb"));
    if (!f.f) return fail(HX3D_EINVAL, std::string("cannot open ") + path);
    char name[NAME_LEN];
    int32_t t, nt, k, it, beg;
    double tm, r[6];
    bool ok = std::fread(name, 1, NAME_LEN, f.f) == (size_t)NAME_LEN && get(f.f, t) && get(f.f, nt) && get(f.f, k) &&
              get(f.f, it) && get(f.f, tm);
    for (int i = 0; i < 6; ++i) ok = ok && get(f.f, r[i]);
    ok = ok && get(f.f, beg);
    if (!ok) return fail(HX3D_EINVAL, std::string("truncated solution file header: ") + path);
    if (beg != BEGINNING_DATA) return fail(HX3D_EINVAL, std::string("not a HORSES3D solution file: ") + path);
    if (type) *type = t;
    if (node_type) *node_type = nt;
    if (K) *K = k;
    if (iter) *iter = it;
    if (time) *time = tm;
    if (refs)
        for (int i = 0; i < 6; ++i) refs[i] = r[i];
    return HX3D_OK;
}

int hx3d_hsol_read(const char *path, int K, const int *N, double *Q, double *grad)
{
    if (!path || !N || !Q || K <= 0) return fail(HX3D_EINVAL, "hx3d_hsol_read: missing argument");
    int type = 0, nt = 0, k = 0;
    int rc = hx3d_hsol_read_header(path, &type, &nt, &k, nullptr, nullptr, nullptr);
    if (rc) return rc;
    // HexMesh_LoadSolution's checks: the element count, and a file type that carries Q
    if (k != K) return fail(HX3D_EINVAL, "solution file has " + std::to_string(k) + " elements, the mesh " + std::to_string(K));
    if (type != SOLUTION_FILE && type != SOLUTION_AND_GRADIENTS_FILE)
        return fail(HX3D_EUNSUPPORTED, "solution file type " + std::to_string(type) + " is not a solution file");
    if (grad && type != SOLUTION_AND_GRADIENTS_FILE) return fail(HX3D_EINVAL, "solution file holds no gradients");
    // a file of the other node type is loaded as it is, without interpolation, as HexMesh_LoadSolution
    // does (HexMesh.f90:4798-4806, with a warning there); the caller sees the type in the header
    if (nt != GAUSSLOBATTO && nt != GAUSS) return fail(HX3D_EUNSUPPORTED, "solution file: unknown node type " + std::to_string(nt));
    long nodes = 0;
    for (int e = 0; e < K; ++e) nodes += (long)(N[3 * e] + 1) * (N[3 * e + 1] + 1) * (N[3 * e + 2] + 1);
    File f(std::fopen(path, "rb"));
    if (!f.f || std::fseek(f.f, POS_INIT_DATA, SEEK_SET) != 0) return fail(HX3D_EINVAL, std::string("cannot read ") + path);
    std::vector<double> skip;
    long off = 0;
    for (int e = 0; e < K; ++e) {
        int32_t d[5];
        for (int i = 0; i < 5; ++i)
            if (!get(f.f, d[i])) return fail(HX3D_EINVAL, "truncated solution file");
        if (d[0] != 4 || d[1] != 5 || d[2] != N[3 * e] + 1 || d[3] != N[3 * e + 1] + 1 || d[4] != N[3 * e + 2] + 1)
            return fail(HX3D_EINVAL, "element " + std::to_string(e + 1) + " of the solution file has other dimensions "
                                     "than the mesh (orders differ)");
        const long ne = (long)d[2] * d[3] * d[4];
        if (std::fread(Q + 5 * off, sizeof(double), 5 * ne, f.f) != (size_t)(5 * ne))
            return fail(HX3D_EINVAL, "truncated solution file");
        if (type == SOLUTION_AND_GRADIENTS_FILE)
            for (int g = 0; g < 3; ++g) {
                double *dst = grad ? grad + 5 * (g * nodes + off) : (skip.resize(5 * ne), skip.data());
                if (std::fread(dst, sizeof(double), 5 * ne, f.f) != (size_t)(5 * ne))
                    return fail(HX3D_EINVAL, "truncated solution file");
            }
        off += ne;
    }
    int32_t eof = 0;
    if (!get(f.f, eof) || eof != END_OF_FILE) return fail(HX3D_EINVAL, "solution file is not sealed (no end marker)");
    return HX3D_OK;
}

}  // extern "C"
