// Mesh front end of the hx3d engine: a HORSES3D SpecMesh file read into the element / face
// contract of hx3d_case_desc, with the reference's geometry.
//
//   file format            Read_SpecMesh.f90:36-348 (ConstructMesh_FromSpecMeshFile_)
//   faces                  ConstructFaces (HexMesh.f90:356-432), faceRotation (:490-515)
//   zones                  ConstructZones (ZoneClass.f90:38-78): names in the order of the
//                          reference's name dictionary (FTDictionaryClass.f90, 8 buckets keyed
//                          by b3hs_hash_key_jenkins, Hash.f90)
//   periodic faces         ConstructPeriodicFaces (HexMesh.f90:519-739),
//                          DeletePeriodicMinusFaces (:872-944), GetElementsFaceIDs (:434-475)
//   face orders            Face_LinkWithElements (FaceClass.f90:187-278)
//   face patches, maps     FacePatchClass.f90, TransfiniteMaps3D.f90:145-881, the patch
//                          projection of HexMesh_ConstructGeometry (HexMesh.f90:2677-2900)
//   element geometry       ConstructMappedGeometry + computeMetricTermsConservativeForm
//                          (MappedGeometry.f90:64-386)
//   face geometry          ConstructMappedGeometryFace (MappedGeometry.f90:482-760), left element
//
// Every arithmetic step keeps the reference's operation order, so that the node coordinates,
// metrics, Jacobians and face frames come out with the reference's bits.  Host code only.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hx3d.h"
#include "frontend.hpp"
#include "spectral.hpp"

struct hx3d_mesh {
    std::string err;
    int K = 0, F = 0, bface = 0;
    int nodes = HX3D_NODES_GAUSSLOBATTO;   // the node type the geometry is built on
    std::vector<int> N, faceID, faceSide, fElem, fESide, fRot, fNf;
    std::vector<long> off, foff;
    std::vector<double> x, Ja, J, fNormal, fT1, fT2, fJac;
};

namespace {

using hxs::almost_equal;

// element face k (0..5 = FRONT, BACK, BOTTOM, RIGHT, TOP, LEFT): its corners (localFaceNode,
// HexElementConnectivityDefinitions.f90:69-77, 0-based) and the element axes its two face
// coordinates run along (axisMap)
constexpr int FACE_NODE[6][4] = {{0, 1, 5, 4}, {3, 2, 6, 7}, {0, 1, 2, 3}, {1, 2, 6, 5}, {4, 5, 6, 7}, {0, 3, 7, 4}};
constexpr int AXIS_MAP[6][2] = {{0, 2}, {0, 2}, {0, 1}, {1, 2}, {0, 1}, {1, 2}};

// ------------------------------------------------------------------------------------
// face patches (FacePatchClass.f90): a polynomial surface through points on tensor knots
// ------------------------------------------------------------------------------------
struct Patch {
    int nu = 0, nv = 0;
    std::vector<double> u, v, wbu, wbv, Du, Dv, p;   // p[3*(i + nu*j) + c]

    void build(const std::vector<double> &uk, const std::vector<double> &vk, const std::vector<double> &pts)
    {
        nu = (int)uk.size();
        nv = (int)vk.size();
        u = uk;
        v = vk;
        p = pts;
        wbu.assign(nu, 0.0);
        wbv.assign(nv, 0.0);
        Du.assign(nu * nu, 0.0);
        Dv.assign(nv * nv, 0.0);
        hxs::bary(nu - 1, u.data(), wbu.data());
        hxs::bary(nv - 1, v.data(), wbv.data());
        hxs::deriv_matrix(nu - 1, u.data(), Du.data());
        hxs::deriv_matrix(nv - 1, v.data(), Dv.data());
    }
    const double *at(int i, int j) const { return &p[3 * (i + nu * j)]; }
    bool quad() const { return nu == 2 && nv == 2; }

    // ComputeFacePoint
    void point(double a, double b, double out[3]) const
    {
        if (quad()) {
            for (int c = 0; c < 3; ++c)
                out[c] = 0.25 * (at(0, 0)[c] * (1.0 - a) * (1.0 - b) + at(1, 0)[c] * (1.0 + a) * (1.0 - b) +
                                 at(1, 1)[c] * (1.0 + a) * (1.0 + b) + at(0, 1)[c] * (1.0 - a) * (1.0 + b));
            return;
        }
        double li[16], lj[16];
        hxs::interp_vector(a, nu - 1, u.data(), wbu.data(), li);
        hxs::interp_vector(b, nv - 1, v.data(), wbv.data(), lj);
        out[0] = out[1] = out[2] = 0.0;
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < nu; ++i)
                for (int c = 0; c < 3; ++c) out[c] = out[c] + at(i, j)[c] * li[i] * lj[j];
    }
    // ComputeFaceDerivative: g[c][0] = d/da, g[c][1] = d/db
    void deriv(double a, double b, double g[3][2]) const
    {
        if (quad()) {
            for (int c = 0; c < 3; ++c) {
                g[c][0] = 0.25 * (-at(0, 0)[c] * (1.0 - b) + at(1, 0)[c] * (1.0 - b) + at(1, 1)[c] * (1.0 + b) -
                                  at(0, 1)[c] * (1.0 + b));
                g[c][1] = 0.25 * (-at(0, 0)[c] * (1.0 - a) - at(1, 0)[c] * (1.0 + a) + at(1, 1)[c] * (1.0 + a) +
                                  at(0, 1)[c] * (1.0 - a));
            }
            return;
        }
        double li[16], lj[16], dli[16], dlj[16];
        hxs::interp_vector(a, nu - 1, u.data(), wbu.data(), li);
        hxs::interp_vector(b, nv - 1, v.data(), wbv.data(), lj);
        for (int i = 0; i < nu; ++i) {
            dli[i] = 0.0;
            for (int j = 0; j < nu; ++j) dli[i] = dli[i] + Du[j * nu + i] * li[j];
        }
        for (int i = 0; i < nv; ++i) {
            dlj[i] = 0.0;
            for (int j = 0; j < nv; ++j) dlj[i] = dlj[i] + Dv[j * nv + i] * lj[j];
        }
        for (int c = 0; c < 3; ++c) g[c][0] = g[c][1] = 0.0;
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < nu; ++i)
                for (int c = 0; c < 3; ++c) {
                    g[c][0] = g[c][0] + at(i, j)[c] * dli[i] * lj[j];
                    g[c][1] = g[c][1] + at(i, j)[c] * li[i] * dlj[j];
                }
    }
    // ProjectFaceToNewPoints onto the CGL points of orders (Mx, My), then rebuilt on them
    void project(int Mx, int My)
    {
        std::vector<double> xu(Mx + 1), xv(My + 1), q(3 * (Mx + 1) * (My + 1));
        hxs::cgl_nodes(Mx, xu.data());
        hxs::cgl_nodes(My, xv.data());
        for (int j = 0; j <= My; ++j)
            for (int i = 0; i <= Mx; ++i) point(xu[i], xv[j], &q[3 * (i + (Mx + 1) * j)]);
        build(xu, xv, q);
    }
};

// ------------------------------------------------------------------------------------
// transfinite hex maps (TransfiniteMaps3D.f90)
// ------------------------------------------------------------------------------------
struct HexMap {
    bool hex8 = true;
    double cn[3][8];   // corners
    Patch f[6];

    // Hex8TransfiniteMap / GeneralHexTransfiniteMap + ComputeHexTransfiniteMap
    void at(const double u[3], double x[3]) const
    {
        double xi[3];
        for (int d = 0; d < 3; ++d) xi[d] = 0.5 * (u[d] + 1.0);
        if (hex8) {
            for (int j = 0; j < 3; ++j)
                x[j] = cn[j][0] * (1.0 - xi[0]) * (1.0 - xi[1]) * (1.0 - xi[2]) +
                       cn[j][1] * xi[0] * (1.0 - xi[1]) * (1.0 - xi[2]) + cn[j][2] * xi[0] * xi[1] * (1.0 - xi[2]) +
                       cn[j][3] * (1.0 - xi[0]) * xi[1] * (1.0 - xi[2]) +
                       cn[j][4] * (1.0 - xi[0]) * (1.0 - xi[1]) * xi[2] + cn[j][5] * xi[0] * (1.0 - xi[1]) * xi[2] +
                       cn[j][6] * xi[0] * xi[1] * xi[2] + cn[j][7] * (1.0 - xi[0]) * xi[1] * xi[2];
            return;
        }
        double E[12][3], P[6][3];
        edges_faces(u, E, P);
        for (int j = 0; j < 3; ++j) {
            double s = P[5][j] * (1.0 - xi[0]) + P[3][j] * xi[0] + P[0][j] * (1.0 - xi[1]) + P[1][j] * xi[1] +
                       P[2][j] * (1.0 - xi[2]) + P[4][j] * xi[2];
            s = s - E[0][j] * (1.0 - xi[1]) * (1.0 - xi[2]) - E[2][j] * (1.0 - xi[1]) * xi[2] -
                E[4][j] * xi[1] * (1.0 - xi[2]) - E[6][j] * xi[1] * xi[2] - E[8][j] * (1.0 - xi[0]) * (1.0 - xi[2]) -
                E[11][j] * (1.0 - xi[0]) * xi[2] - E[9][j] * (1.0 - xi[2]) * xi[0] - E[10][j] * xi[0] * xi[2] -
                E[3][j] * (1.0 - xi[0]) * (1.0 - xi[1]) - E[7][j] * (1.0 - xi[0]) * xi[1] -
                E[1][j] * xi[0] * (1.0 - xi[1]) - E[5][j] * xi[0] * xi[1];
            s = s + cn[j][0] * (1.0 - xi[0]) * (1.0 - xi[1]) * (1.0 - xi[2]) +
                cn[j][4] * (1.0 - xi[0]) * (1.0 - xi[1]) * xi[2] + cn[j][3] * (1.0 - xi[0]) * xi[1] * (1.0 - xi[2]) +
                cn[j][7] * (1.0 - xi[0]) * xi[1] * xi[2] + cn[j][1] * xi[0] * (1.0 - xi[1]) * (1.0 - xi[2]) +
                cn[j][5] * xi[0] * (1.0 - xi[1]) * xi[2] + cn[j][2] * xi[0] * xi[1] * (1.0 - xi[2]) +
                cn[j][6] * xi[0] * xi[1] * xi[2];
            x[j] = s;
        }
    }

    // the 12 edge points and 6 face points of GeneralHexTransfiniteMap
    void edges_faces(const double u[3], double E[12][3], double P[6][3]) const
    {
        f[0].point(u[0], -1.0, E[0]);
        f[0].point(1.0, u[2], E[1]);
        f[0].point(u[0], 1.0, E[2]);
        f[0].point(-1.0, u[2], E[3]);
        f[1].point(u[0], -1.0, E[4]);
        f[1].point(1.0, u[2], E[5]);
        f[1].point(u[0], 1.0, E[6]);
        f[1].point(-1.0, u[2], E[7]);
        f[3].point(u[1], -1.0, E[9]);
        f[5].point(u[1], -1.0, E[8]);
        f[3].point(u[1], 1.0, E[10]);
        f[5].point(u[1], 1.0, E[11]);
        f[0].point(u[0], u[2], P[0]);
        f[1].point(u[0], u[2], P[1]);
        f[2].point(u[0], u[1], P[2]);
        f[3].point(u[1], u[2], P[3]);
        f[4].point(u[0], u[1], P[4]);
        f[5].point(u[1], u[2], P[5]);
    }

    // GradHex8TransfiniteMap / GradGeneralHexTransfiniteMap + ComputeGradHexTransfiniteMap:
    // g[c][m] = d x_c / d u_m
    void grad(const double u[3], double g[3][3]) const
    {
        double xi[3];
        for (int d = 0; d < 3; ++d) xi[d] = 0.5 * (u[d] + 1.0);
        if (hex8) {
            for (int i = 0; i < 3; ++i) {
                g[i][0] = -cn[i][0] * (1.0 - xi[1]) * (1.0 - xi[2]) + cn[i][1] * (1.0 - xi[1]) * (1.0 - xi[2]) +
                          cn[i][2] * xi[1] * (1.0 - xi[2]) - cn[i][3] * xi[1] * (1.0 - xi[2]) -
                          cn[i][4] * (1.0 - xi[1]) * xi[2] + cn[i][5] * (1.0 - xi[1]) * xi[2] +
                          cn[i][6] * xi[1] * xi[2] - cn[i][7] * xi[1] * xi[2];
                g[i][1] = -cn[i][0] * (1.0 - xi[0]) * (1.0 - xi[2]) - cn[i][1] * xi[0] * (1.0 - xi[2]) +
                          cn[i][2] * xi[0] * (1.0 - xi[2]) + cn[i][3] * (1.0 - xi[0]) * (1.0 - xi[2]) -
                          cn[i][4] * (1.0 - xi[0]) * xi[2] - cn[i][5] * xi[0] * xi[2] + cn[i][6] * xi[0] * xi[2] +
                          cn[i][7] * (1.0 - xi[0]) * xi[2];
                g[i][2] = -cn[i][0] * (1.0 - xi[0]) * (1.0 - xi[1]) - cn[i][1] * xi[0] * (1.0 - xi[1]) -
                          cn[i][2] * xi[0] * xi[1] - cn[i][3] * (1.0 - xi[0]) * xi[1] +
                          cn[i][4] * (1.0 - xi[0]) * (1.0 - xi[1]) + cn[i][5] * xi[0] * (1.0 - xi[1]) +
                          cn[i][6] * xi[0] * xi[1] + cn[i][7] * (1.0 - xi[0]) * xi[1];
            }
            for (int i = 0; i < 3; ++i)
                for (int m = 0; m < 3; ++m) g[i][m] = 0.5 * g[i][m];
            return;
        }
        double E[12][3], P[6][3], ED[12][3], FD[6][3][2], t[3][2];
        edges_faces(u, E, P);
        auto take = [&](int k, int face, double a, double b, int which) {
            f[face].deriv(a, b, t);
            for (int c = 0; c < 3; ++c) ED[k][c] = 2.0 * t[c][which];
        };
        take(0, 0, u[0], -1.0, 0);
        take(1, 0, 1.0, u[2], 1);
        take(2, 0, u[0], 1.0, 0);
        take(3, 0, -1.0, u[2], 1);
        take(4, 1, u[0], -1.0, 0);
        take(5, 1, 1.0, u[2], 1);
        take(6, 1, u[0], 1.0, 0);
        take(7, 1, -1.0, u[2], 1);
        take(8, 5, u[1], -1.0, 0);
        take(9, 3, u[1], -1.0, 0);
        take(10, 3, u[1], 1.0, 0);
        take(11, 5, u[1], 1.0, 0);
        const double fa[6][2] = {{u[0], u[2]}, {u[0], u[2]}, {u[0], u[1]}, {u[1], u[2]}, {u[0], u[1]}, {u[1], u[2]}};
        for (int k = 0; k < 6; ++k) {
            f[k].deriv(fa[k][0], fa[k][1], t);
            for (int c = 0; c < 3; ++c) { FD[k][c][0] = 2.0 * t[c][0]; FD[k][c][1] = 2.0 * t[c][1]; }
        }
        for (int j = 0; j < 3; ++j) {
            double s = -P[5][j] + P[3][j] + FD[0][j][0] * (1.0 - xi[1]) + FD[1][j][0] * xi[1] +
                       FD[2][j][0] * (1.0 - xi[2]) + FD[4][j][0] * xi[2];
            s = s - ED[0][j] * (1.0 - xi[1]) * (1.0 - xi[2]) - ED[2][j] * (1.0 - xi[1]) * xi[2] -
                ED[4][j] * xi[1] * (1.0 - xi[2]) - ED[6][j] * xi[1] * xi[2] + E[8][j] * (1.0 - xi[2]) +
                E[11][j] * xi[2] - E[9][j] * (1.0 - xi[2]) - E[10][j] * xi[2] + E[3][j] * (1.0 - xi[1]) +
                E[7][j] * xi[1] - E[1][j] * (1.0 - xi[1]) - E[5][j] * xi[1];
            s = s - cn[j][0] * (1.0 - xi[1]) * (1.0 - xi[2]) - cn[j][4] * (1.0 - xi[1]) * xi[2] -
                cn[j][3] * xi[1] * (1.0 - xi[2]) - cn[j][7] * xi[1] * xi[2] + cn[j][1] * (1.0 - xi[1]) * (1.0 - xi[2]) +
                cn[j][5] * (1.0 - xi[1]) * xi[2] + cn[j][2] * xi[1] * (1.0 - xi[2]) + cn[j][6] * xi[1] * xi[2];
            g[j][0] = s;

            s = FD[5][j][0] * (1.0 - xi[0]) + FD[3][j][0] * xi[0] - P[0][j] + P[1][j] + FD[2][j][1] * (1.0 - xi[2]) +
                FD[4][j][1] * xi[2];
            s = s + E[0][j] * (1.0 - xi[2]) + E[2][j] * xi[2] - E[4][j] * (1.0 - xi[2]) - E[6][j] * xi[2] -
                ED[8][j] * (1.0 - xi[0]) * (1.0 - xi[2]) - ED[11][j] * (1.0 - xi[0]) * xi[2] -
                ED[9][j] * (1.0 - xi[2]) * xi[0] - ED[10][j] * xi[0] * xi[2] + E[3][j] * (1.0 - xi[0]) -
                E[7][j] * (1.0 - xi[0]) + E[1][j] * xi[0] - E[5][j] * xi[0];
            s = s - cn[j][0] * (1.0 - xi[0]) * (1.0 - xi[2]) - cn[j][4] * (1.0 - xi[0]) * xi[2] +
                cn[j][3] * (1.0 - xi[0]) * (1.0 - xi[2]) + cn[j][7] * (1.0 - xi[0]) * xi[2] -
                cn[j][1] * xi[0] * (1.0 - xi[2]) - cn[j][5] * xi[0] * xi[2] + cn[j][2] * xi[0] * (1.0 - xi[2]) +
                cn[j][6] * xi[0] * xi[2];
            g[j][1] = s;

            s = FD[5][j][1] * (1.0 - xi[0]) + FD[3][j][1] * xi[0] + FD[0][j][1] * (1.0 - xi[1]) +
                FD[1][j][1] * xi[1] - P[2][j] + P[4][j];
            s = s + E[0][j] * (1.0 - xi[1]) - E[2][j] * (1.0 - xi[1]) + E[4][j] * xi[1] - E[6][j] * xi[1] +
                E[8][j] * (1.0 - xi[0]) - E[11][j] * (1.0 - xi[0]) + E[9][j] * xi[0] - E[10][j] * xi[0] -
                ED[3][j] * (1.0 - xi[0]) * (1.0 - xi[1]) - ED[7][j] * (1.0 - xi[0]) * xi[1] -
                ED[1][j] * xi[0] * (1.0 - xi[1]) - ED[5][j] * xi[0] * xi[1];
            s = s - cn[j][0] * (1.0 - xi[0]) * (1.0 - xi[1]) + cn[j][4] * (1.0 - xi[0]) * (1.0 - xi[1]) -
                cn[j][3] * (1.0 - xi[0]) * xi[1] + cn[j][7] * (1.0 - xi[0]) * xi[1] -
                cn[j][1] * xi[0] * (1.0 - xi[1]) + cn[j][5] * xi[0] * (1.0 - xi[1]) - cn[j][2] * xi[0] * xi[1] +
                cn[j][6] * xi[0] * xi[1];
            g[j][2] = s;
        }
        const double eps = 100.0 * 2.220446049250313e-16;
        for (int i = 0; i < 3; ++i)
            for (int m = 0; m < 3; ++m) {
                g[i][m] = 0.5 * g[i][m];
                if (std::fabs(g[i][m]) <= eps) g[i][m] = 0.0;
            }
    }
};

// b3hs_hash_key_jenkins (Hash.f90): Bob Jenkins' 1996 hash ("lookup2") computed, as the
// reference does, in 64-bit integers; the bucket is the low bits of c, 1-based
int jenkins_bucket(const std::string &key, int range)
{
    uint64_t a = (uint64_t)(int64_t)-1640531527, b = a, c = 305419896;
    auto mix = [&]() {
        a = (a - b - c) ^ (c >> 13);
        b = (b - c - a) ^ (a << 8);
        c = (c - a - b) ^ (b >> 13);
        a = (a - b - c) ^ (c >> 12);
        b = (b - c - a) ^ (a << 16);
        c = (c - a - b) ^ (b >> 5);
        a = (a - b - c) ^ (c >> 3);
        b = (b - c - a) ^ (a << 10);
        c = (c - a - b) ^ (b >> 15);
    };
    const unsigned char *k = (const unsigned char *)key.data();
    size_t len = key.size();
    auto ch = [&](size_t i, int sh) { return (uint64_t)k[i] << sh; };
    while (len >= 12) {
        a += ch(0, 0) + ch(1, 8) + ch(2, 16) + ch(3, 24);
        b += ch(4, 0) + ch(5, 8) + ch(6, 16) + ch(7, 24);
        c += ch(8, 0) + ch(9, 8) + ch(10, 16) + ch(11, 24);
        mix();
        k += 12;
        len -= 12;
    }
    c += len;
    if (len >= 11) c += ch(10, 24);
    if (len >= 10) c += ch(9, 16);
    if (len >= 9) c += ch(8, 8);
    if (len >= 8) b += ch(7, 24);
    if (len >= 7) b += ch(6, 16);
    if (len >= 6) b += ch(5, 8);
    if (len >= 5) b += ch(4, 0);
    if (len >= 4) a += ch(3, 24);
    if (len >= 3) a += ch(2, 16);
    if (len >= 2) a += ch(1, 8);
    if (len >= 1) a += ch(0, 0);
    mix();
    return (int)(c & (uint64_t)(range - 1)) + 1;
}

// NORM2 of a 3-vector as the reference build evaluates it (the running-maximum scaled sum of
// the Fortran runtime: m sqrt(1 + sum((x_i / m)^2)) over the other entries), checked bit for bit
// against the intrinsic on 20,000 random vectors; used for the face Jacobian and tangent norms
double norm2(const double v[3])
{
    double mx = 0.0, s = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double a = std::fabs(v[i]);
        if (mx == 0.0) {
            mx = a;
        } else if (a > mx) {
            const double t = mx / a, tsq = t * t;
            s = s * tsq;
            s = s + tsq;
            mx = a;
        } else {
            const double t = a / mx;
            s = s + t * t;
        }
    }
    return mx * std::sqrt(1.0 + s);
}

// faceRotation (HexMesh.f90:490-515)
int face_rotation(const int m[4], const int s[4])
{
    static const int next[4] = {1, 2, 3, 0};
    int j = 0;
    for (; j < 4; ++j)
        if (m[0] == s[j]) break;
    if (j == 4) j = 3;   // as the reference's loop leaves it when nothing matches
    return m[1] == s[next[j]] ? j : j + 4;
}

// CompareTwoNodes (HexMesh.f90:741-800): coord = 0 finds the one differing coordinate
bool compare_nodes(const double *x1, const double *x2, int &coord)
{
    int counter = 0;
    if (coord == 0) {
        int c = 0;
        for (int i = 0; i < 3; ++i) {
            if (almost_equal(x1[i], x2[i])) ++counter;
            else c = i + 1;
        }
        if (c) coord = c;
        return counter >= 2;
    }
    for (int i = 0; i < 3; ++i)
        if (i + 1 != coord && almost_equal(x1[i], x2[i])) ++counter;
    return counter >= 2;
}

std::string lower(std::string s)
{
    for (auto &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

struct FaceT {
    int nodes[4];
    int elem[2] = {-1, -1}, side[2] = {0, 0};   // side 1..6
    int rot = 0;
    bool interior = false;
    std::string bname;
    int zone = -1;
};

int fail(hx3d_mesh *m, int code, const std::string &msg)
{
    m->err = msg;
    return code;
}

int read_specmesh(hx3d_mesh *m, const char *path, int Nuni, const int *orders, long norders,
                  const hx3d_front::ZoneResolver &resolve, bool gauss)
{
    m->nodes = gauss ? HX3D_NODES_GAUSS : HX3D_NODES_GAUSSLOBATTO;
    std::ifstream in(path);
    if (!in) return fail(m, HX3D_EINVAL, std::string("cannot open mesh file ") + path);
    long nn = 0, ne = 0;
    int bf = 0;
    if (!(in >> nn >> ne >> bf) || nn <= 0 || ne <= 0 || bf < 1)
        return fail(m, HX3D_EINVAL, "bad SpecMesh header (nodes, elements, boundary face order)");
    std::vector<double> X(3 * nn);
    for (long j = 0; j < nn; ++j)
        if (!(in >> X[3 * j] >> X[3 * j + 1] >> X[3 * j + 2])) return fail(m, HX3D_EINVAL, "truncated node list");
    const int K = (int)ne;
    if (orders && norders >= 0 && norders != ne)
        return fail(m, HX3D_EINVAL, "the polynomial orders are given for " + std::to_string(norders) +
                                        " elements, the mesh has " + std::to_string(ne));
    m->K = K;
    m->bface = bf;
    m->N.assign(3 * K, 0);
    for (int e = 0; e < K; ++e)
        for (int d = 0; d < 3; ++d) {
            const int o = !orders ? Nuni : orders[norders == hx3d_front::ORDERS_ONE_ROW ? d : 3 * e + d];
            if (o < 1 || o > 7) return fail(m, HX3D_EUNSUPPORTED, "orders must be 1..7");
            m->N[3 * e + d] = o;
        }
    std::vector<double> cgl(bf + 1);
    hxs::cgl_nodes(bf, cgl.data());
    std::vector<std::array<int, 8>> enodes(K);
    std::vector<HexMap> maps(K);
    std::vector<std::array<std::string, 6>> bnames(K);
    std::vector<std::string> zone_insert;   // dictionary keys in insertion order
    std::set<std::string> zone_seen;
    const std::vector<double> flat = {-1.0, 1.0};
    for (int e = 0; e < K; ++e) {
        int ids[8], flags[6];
        for (int c = 0; c < 8; ++c)
            if (!(in >> ids[c]) || ids[c] < 1 || ids[c] > nn) return fail(m, HX3D_EINVAL, "bad element node list");
        for (int k = 0; k < 6; ++k)
            if (!(in >> flags[k])) return fail(m, HX3D_EINVAL, "bad element face flags");
        for (int c = 0; c < 8; ++c) enodes[e][c] = ids[c] - 1;
        HexMap &hm = maps[e];
        int fmax = 0;
        for (int k = 0; k < 6; ++k) fmax = std::max(fmax, flags[k]);
        if (fmax == 0) {
            hm.hex8 = true;
            for (int c = 0; c < 8; ++c)
                for (int j = 0; j < 3; ++j) hm.cn[j][c] = X[3 * enodes[e][c] + j];
            // straight faces as bilinear patches (the reference keeps them for IsHex8 elements too)
            for (int k = 0; k < 6; ++k) {
                std::vector<double> q(12);
                const int map4[4] = {0, 1, 3, 2};   // (1,1), (2,1), (2,2), (1,2) <- face corners 1..4
                for (int t = 0; t < 4; ++t)
                    for (int j = 0; j < 3; ++j) q[3 * map4[t] + j] = X[3 * enodes[e][FACE_NODE[k][t]] + j];
                hm.f[k].build(flat, flat, q);
            }
        } else {
            hm.hex8 = false;
            for (int k = 0; k < 6; ++k) {
                if (flags[k] == 0) {
                    std::vector<double> q(12);
                    const int map4[4] = {0, 1, 3, 2};
                    for (int t = 0; t < 4; ++t)
                        for (int j = 0; j < 3; ++j) q[3 * map4[t] + j] = X[3 * enodes[e][FACE_NODE[k][t]] + j];
                    hm.f[k].build(flat, flat, q);
                } else {
                    std::vector<double> q(3 * (bf + 1) * (bf + 1));
                    for (int j = 0; j <= bf; ++j)
                        for (int i = 0; i <= bf; ++i) {
                            double *p = &q[3 * (i + (bf + 1) * j)];
                            if (!(in >> p[0] >> p[1] >> p[2])) return fail(m, HX3D_EINVAL, "truncated face patch");
                        }
                    hm.f[k].build(cgl, cgl, q);
                }
            }
        }
        for (int k = 0; k < 6; ++k) {
            if (!(in >> bnames[e][k])) return fail(m, HX3D_EINVAL, "bad boundary name line");
            if (bnames[e][k] != "---" && !zone_seen.count(bnames[e][k])) {
                zone_seen.insert(bnames[e][k]);
                zone_insert.push_back(bnames[e][k]);
            }
        }
    }
    // ---- faces (ConstructFaces): created in element / face order, matched by node set
    std::vector<FaceT> faces;
    std::map<std::array<int, 4>, int> table;
    for (int e = 0; e < K; ++e)
        for (int k = 0; k < 6; ++k) {
            int fn[4];
            for (int t = 0; t < 4; ++t) fn[t] = enodes[e][FACE_NODE[k][t]];
            std::array<int, 4> key = {fn[0], fn[1], fn[2], fn[3]};
            std::sort(key.begin(), key.end());
            auto it = table.find(key);
            if (it != table.end()) {
                FaceT &f = faces[it->second];
                if (f.interior) return fail(m, HX3D_EINVAL, "a face is shared by more than two elements");
                f.elem[1] = e;
                f.side[1] = k + 1;
                f.interior = true;
                f.rot = face_rotation(f.nodes, fn);
            } else {
                FaceT f;
                std::memcpy(f.nodes, fn, sizeof fn);
                f.elem[0] = e;
                f.side[0] = k + 1;
                f.bname = bnames[e][k];
                table[key] = (int)faces.size();
                faces.push_back(f);
            }
        }
    // ---- zones in the order of the reference's name dictionary (8 buckets, each in insertion order)
    std::vector<std::string> zones;
    for (int bkt = 1; bkt <= 8; ++bkt)
        for (auto &nm : zone_insert)
            if (jenkins_bucket(nm, 8) == bkt) zones.push_back(lower(nm));
    std::vector<std::vector<int>> zfaces(zones.size());
    for (size_t f = 0; f < faces.size(); ++f) {
        if (faces[f].interior) continue;
        const std::string b = lower(faces[f].bname);
        for (size_t z = 0; z < zones.size(); ++z)
            if (zones[z] == b) { faces[f].zone = (int)z; zfaces[z].push_back((int)f); break; }
    }
    // boundary conditions of the zones: periodic ones carry their coupled zone
    std::map<std::string, std::string> coupled;
    for (const auto &z : zones) {
        std::string c, err;
        const int r = resolve(z, c, err);
        if (r < 0) return fail(m, r, err);
        if (r == 1) coupled[z] = c;
    }
    // ---- ConstructPeriodicFaces
    std::vector<char> deleted(zones.size(), 0);
    for (size_t zp = 0; zp < zones.size(); ++zp) {
        auto cit = coupled.find(zones[zp]);
        if (cit == coupled.end() || deleted[zp]) continue;
        size_t zm = zones.size();
        for (size_t z = 0; z < zones.size(); ++z)
            if (zones[z] == cit->second) { zm = z; break; }
        if (zm == zones.size())
            return fail(m, HX3D_EINVAL, "coupled boundary \"" + cit->second + "\" of \"" + zones[zp] + "\" not found");
        deleted[zm] = 1;
        int coord = 0;
        for (int fi : zfaces[zp]) {
            FaceT &F = faces[fi];
            if (F.interior) continue;
            bool found = false;
            for (int fj : zfaces[zm]) {
                FaceT &G = faces[fj];
                if (G.interior) continue;
                bool mm[4] = {false, false, false, false};
                if (coord == 0) {
                    for (int lc = 1; lc <= 3; ++lc) {
                        bool sm[4] = {false, false, false, false};
                        for (int t = 0; t < 4; ++t) mm[t] = false;
                        for (int k = 0; k < 4; ++k) {
                            for (int l = 0; l < 4; ++l) {
                                if (sm[l]) continue;
                                int c = lc;
                                if (compare_nodes(&X[3 * F.nodes[k]], &X[3 * G.nodes[l]], c)) {
                                    mm[k] = true;
                                    sm[l] = true;
                                    break;
                                }
                            }
                            if (!mm[k]) break;
                        }
                        if (mm[0] && mm[1] && mm[2] && mm[3]) { coord = lc; break; }
                    }
                } else {
                    bool sm[4] = {false, false, false, false};
                    for (int k = 0; k < 4; ++k) {
                        for (int l = 0; l < 4; ++l) {
                            if (sm[l]) continue;
                            int c = coord;
                            if (compare_nodes(&X[3 * F.nodes[k]], &X[3 * G.nodes[l]], c)) {
                                mm[k] = true;
                                sm[l] = true;
                                break;
                            }
                        }
                        if (!mm[k]) break;
                    }
                }
                if (!(mm[0] && mm[1] && mm[2] && mm[3])) continue;
                F.bname.clear();
                F.elem[1] = G.elem[0];
                F.side[1] = G.side[0];
                F.interior = true;
                int slave[4] = {-1, -1, -1, -1};
                for (int k = 0; k < 4; ++k)
                    for (int l = 0; l < 4; ++l) {
                        int c = coord;
                        if (compare_nodes(&X[3 * F.nodes[k]], &X[3 * G.nodes[l]], c)) slave[l] = F.nodes[k];
                    }
                F.rot = face_rotation(F.nodes, slave);
                found = true;
                break;
            }
            if (!found)
                return fail(m, HX3D_EINVAL, "a face of periodic boundary \"" + zones[zp] + "\" has no partner");
        }
    }
    // ---- DeletePeriodicMinusFaces, then every remaining face must be interior
    std::vector<FaceT> kept;
    for (auto &f : faces) {
        if (f.interior) { kept.push_back(f); continue; }
        if (f.zone >= 0 && deleted[f.zone]) continue;
        return fail(m, HX3D_EUNSUPPORTED, "boundary \"" + f.bname + "\" is not periodic: the engine takes meshes whose "
                                          "faces are all interior or periodic");
    }
    const int F = (int)kept.size();
    m->F = F;
    m->faceID.assign(6 * K, -1);
    m->faceSide.assign(6 * K, 0);
    m->fElem.resize(2 * F);
    m->fESide.resize(2 * F);
    m->fRot.resize(F);
    for (int f = 0; f < F; ++f) {
        for (int s = 0; s < 2; ++s) {
            m->fElem[2 * f + s] = kept[f].elem[s];
            m->fESide[2 * f + s] = kept[f].side[s];
            m->faceID[6 * kept[f].elem[s] + kept[f].side[s] - 1] = f;
            m->faceSide[6 * kept[f].elem[s] + kept[f].side[s] - 1] = s + 1;
        }
        m->fRot[f] = kept[f].rot;
    }
    // ---- face orders (Face_LinkWithElements): each side's orders along the face, the right
    // side's in the left frame (swapped by rotations 1, 3, 4, 6), the face order their maximum
    m->fNf.resize(2 * F);
    std::vector<int> NfL(2 * F), NfR(2 * F);
    for (int f = 0; f < F; ++f) {
        const int eL = kept[f].elem[0], sL = kept[f].side[0] - 1, eR = kept[f].elem[1], sR = kept[f].side[1] - 1;
        int l[2] = {m->N[3 * eL + AXIS_MAP[sL][0]], m->N[3 * eL + AXIS_MAP[sL][1]]};
        int r[2] = {m->N[3 * eR + AXIS_MAP[sR][0]], m->N[3 * eR + AXIS_MAP[sR][1]]};
        const int rot = kept[f].rot;
        if (rot == 1 || rot == 3 || rot == 4 || rot == 6) std::swap(r[0], r[1]);
        NfL[2 * f] = l[0];
        NfL[2 * f + 1] = l[1];
        NfR[2 * f] = r[0];
        NfR[2 * f + 1] = r[1];
        m->fNf[2 * f] = std::max(l[0], r[0]);
        m->fNf[2 * f + 1] = std::max(l[1], r[1]);
    }
    // ---- face patch projection of HexMesh_ConstructGeometry
    for (int f = 0; f < F; ++f) {
        const int eL = kept[f].elem[0], sL = kept[f].side[0] - 1, eR = kept[f].elem[1], sR = kept[f].side[1] - 1;
        Patch &pL = maps[eL].f[sL], &pR = maps[eR].f[sR];
        const int nsL[2] = {pL.nu - 1, pL.nv - 1}, nsR[2] = {pR.nu - 1, pR.nv - 1};
        const bool h8L = maps[eL].hex8, h8R = maps[eR].hex8;
        if (h8L && h8R) continue;
        if (h8L && nsR[0] == 1 && nsR[1] == 1) continue;
        if (h8R && nsL[0] == 1 && nsL[1] == 1) continue;
        if (nsL[0] == 1 && nsL[1] == 1 && nsR[0] == 1 && nsR[1] == 1) continue;
        int cln[2] = {std::min(NfL[2 * f], NfR[2 * f]), std::min(NfL[2 * f + 1], NfR[2 * f + 1])};
        if (cln[0] < nsL[0] || cln[1] < nsL[1]) pL.project(cln[0], cln[1]);
        const int rot = kept[f].rot;
        if ((rot == 1 || rot == 3 || rot == 4 || rot == 6) && cln[0] != cln[1]) std::swap(cln[0], cln[1]);
        if (cln[0] < nsR[0] || cln[1] < nsR[1]) pR.project(cln[0], cln[1]);
    }
    // corners of the general maps from the (projected) bottom and top patches (constructWithFaces)
    for (int e = 0; e < K; ++e) {
        HexMap &hm = maps[e];
        if (hm.hex8) continue;
        const double uv[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        for (int c = 0; c < 4; ++c) {
            double p[3];
            hm.f[2].point(uv[c][0], uv[c][1], p);
            for (int j = 0; j < 3; ++j) hm.cn[j][c] = p[j];
            hm.f[4].point(uv[c][0], uv[c][1], p);
            for (int j = 0; j < 3; ++j) hm.cn[j][c + 4] = p[j];
        }
    }
    // ---- element geometry (ConstructMappedGeometry, conservative metrics)
    m->off.resize(K + 1);
    m->off[0] = 0;
    for (int e = 0; e < K; ++e)
        m->off[e + 1] = m->off[e] + (long)(m->N[3 * e] + 1) * (m->N[3 * e + 1] + 1) * (m->N[3 * e + 2] + 1);
    const long nodes = m->off[K];
    m->x.assign(3 * nodes, 0.0);
    m->Ja.assign(9 * nodes, 0.0);
    m->J.assign(nodes, 0.0);
    double xg[8][8], wg[8][8], xc[8][8], Dc[8][64], Tc[8][64], wbc[8][8], Dg[8][64];
    for (int n = 1; n <= 7; ++n) {
        hxs::nodes_of(gauss, n, xg[n], wg[n]);
        hxs::cgl_nodes(n, xc[n]);
        hxs::bary(n, xc[n], wbc[n]);
        hxs::deriv_matrix(n, xc[n], Dc[n]);
        hxs::interp_matrix(n, n, xc[n], wbc[n], xg[n], Tc[n]);   // TCheb2Gauss
        hxs::deriv_matrix(n, xg[n], Dg[n]);
    }
    std::vector<double> invJ(nodes);
    for (int e = 0; e < K; ++e) {
        const HexMap &hm = maps[e];
        const int Nx = m->N[3 * e], Ny = m->N[3 * e + 1], Nz = m->N[3 * e + 2];
        const int nx = Nx + 1, ny = Ny + 1, nz = Nz + 1, ne = nx * ny * nz;
        auto id = [&](int i, int j, int k) { return i + nx * (j + ny * k); };
        const long o = m->off[e];
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double u[3] = {xg[Nx][i], xg[Ny][j], xg[Nz][k]};
                    hm.at(u, &m->x[3 * (o + id(i, j, k))]);
                }
        std::vector<double> xC(3 * ne), gC(9 * ne), JaC(9 * ne), JC(ne), cp(3 * ne), ag(9 * ne);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double u[3] = {xc[Nx][i], xc[Ny][j], xc[Nz][k]};
                    const int n = id(i, j, k);
                    hm.at(u, &xC[3 * n]);
                    double g[3][3];
                    hm.grad(u, g);
                    for (int c = 0; c < 3; ++c)
                        for (int mm = 0; mm < 3; ++mm) gC[9 * n + 3 * c + mm] = g[c][mm];
                }
        // coordsProduct = x_a grad x_b - x_b grad x_a for (a, b) = (3, 2), (1, 3), (2, 1), the
        // CGL gradient of it, Ja^d_comp = -1/2 (curl)_d
        const int ab[3][2] = {{2, 1}, {0, 2}, {1, 0}};
        for (int comp = 0; comp < 3; ++comp) {
            const int a = ab[comp][0], b = ab[comp][1];
            for (int n = 0; n < ne; ++n)
                for (int mm = 0; mm < 3; ++mm)
                    cp[3 * n + mm] = xC[3 * n + a] * gC[9 * n + 3 * b + mm] - xC[3 * n + b] * gC[9 * n + 3 * a + mm];
            std::fill(ag.begin(), ag.end(), 0.0);
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        double *A = &ag[9 * id(i, j, k)];   // A[3*mm + l]: d cp_mm / d xi_l
                        for (int l = 0; l < nx; ++l)
                            for (int mm = 0; mm < 3; ++mm)
                                A[3 * mm + 0] = A[3 * mm + 0] + cp[3 * id(l, j, k) + mm] * Dc[Nx][i * nx + l];
                        for (int l = 0; l < ny; ++l)
                            for (int mm = 0; mm < 3; ++mm)
                                A[3 * mm + 1] = A[3 * mm + 1] + cp[3 * id(i, l, k) + mm] * Dc[Ny][j * ny + l];
                        for (int l = 0; l < nz; ++l)
                            for (int mm = 0; mm < 3; ++mm)
                                A[3 * mm + 2] = A[3 * mm + 2] + cp[3 * id(i, j, l) + mm] * Dc[Nz][k * nz + l];
                    }
            for (int n = 0; n < ne; ++n) {
                const double *A = &ag[9 * n];
                const double j1 = A[3 * 2 + 1] - A[3 * 1 + 2], j2 = A[3 * 0 + 2] - A[3 * 2 + 0],
                             j3 = A[3 * 1 + 0] - A[3 * 0 + 1];
                JaC[9 * n + 3 * 0 + comp] = -0.5 * j1;
                JaC[9 * n + 3 * 1 + comp] = -0.5 * j2;
                JaC[9 * n + 3 * 2 + comp] = -0.5 * j3;
            }
        }
        for (int n = 0; n < ne; ++n) {
            // jacobian3D = a1 . (a2 x a3)
            const double *g = &gC[9 * n];
            const double a1[3] = {g[0], g[3], g[6]}, a2[3] = {g[1], g[4], g[7]}, a3[3] = {g[2], g[5], g[8]};
            const double v[3] = {a2[1] * a3[2] - a2[2] * a3[1], a2[2] * a3[0] - a2[0] * a3[2], a2[0] * a3[1] - a2[1] * a3[0]};
            JC[n] = a1[0] * v[0] + a1[1] * v[1] + a1[2] * v[2];
        }
        // TCheb2Gauss in all three directions, one product per term
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const long g = o + id(i, j, k);
                    double ja[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, jj = 0.0;
                    for (int nn2 = 0; nn2 < nz; ++nn2)
                        for (int mm = 0; mm < ny; ++mm)
                            for (int l = 0; l < nx; ++l) {
                                const int s = id(l, mm, nn2);
                                const double tx = Tc[Nx][i * nx + l], ty = Tc[Ny][j * ny + mm], tz = Tc[Nz][k * nz + nn2];
                                for (int d = 0; d < 3; ++d)
                                    for (int c = 0; c < 3; ++c)
                                        ja[3 * d + c] = ja[3 * d + c] + JaC[9 * s + 3 * d + c] * tx * ty * tz;
                                jj = jj + JC[s] * tx * ty * tz;
                            }
                    for (int c = 0; c < 9; ++c) m->Ja[9 * g + c] = ja[c];
                    m->J[g] = jj;
                    invJ[g] = 1.0 / jj;
                }
    }
    // ---- face geometry from the left element (ConstructMappedGeometryFace, eSide 1, rot 0)
    m->foff.resize(F + 1);
    m->foff[0] = 0;
    for (int f = 0; f < F; ++f) m->foff[f + 1] = m->foff[f] + (long)(m->fNf[2 * f] + 1) * (m->fNf[2 * f + 1] + 1);
    const long fn = m->foff[F];
    m->fNormal.assign(3 * fn, 0.0);
    m->fT1.assign(3 * fn, 0.0);
    m->fT2.assign(3 * fn, 0.0);
    m->fJac.assign(fn, 0.0);
    for (int f = 0; f < F; ++f) {
        const int e = kept[f].elem[0], side = kept[f].side[0];   // 1..6
        const int nx = m->N[3 * e] + 1, ny = m->N[3 * e + 1] + 1;
        const int Nf1 = m->fNf[2 * f], Nf2 = m->fNf[2 * f + 1];
        const int Ne1 = NfL[2 * f], Ne2 = NfL[2 * f + 1];   // the element's face orders
        const long o = m->off[e];
        auto id = [&](int i, int j, int k) { return o + i + nx * (j + ny * k); };
        const int fx = Nf1 + 1, fy = Nf2 + 1, ex = Ne1 + 1, ey = Ne2 + 1;
        std::vector<double> xf(3 * fx * fy), dS(3 * ex * ey), nrm(3 * fx * fy);
        // face points through the map (coordRotation with rot = 0 is the identity)
        for (int j = 0; j < fy; ++j)
            for (int i = 0; i < fx; ++i) {
                const double a = xg[Nf1][i], b = xg[Nf2][j];
                double u[3];
                switch (side) {
                case 6: u[0] = -1.0; u[1] = a; u[2] = b; break;
                case 4: u[0] = 1.0; u[1] = a; u[2] = b; break;
                case 3: u[0] = a; u[1] = b; u[2] = -1.0; break;
                case 5: u[0] = a; u[1] = b; u[2] = 1.0; break;
                case 1: u[0] = a; u[1] = -1.0; u[2] = b; break;
                default: u[0] = a; u[1] = 1.0; u[2] = b; break;
                }
                maps[e].at(u, &xf[3 * (i + fx * j)]);
            }
        // dS: the element's metric row of the face's normal direction traced to the face with
        // the Lagrange basis at -1 / +1 (for Gauss-Lobatto nodes the end node itself)
        const int d0 = (side == 6 || side == 4) ? 0 : (side == 3 || side == 5) ? 2 : 1;
        const bool minus = side == 6 || side == 3 || side == 1;
        const int Nd = m->N[3 * e + d0];
        double vb[8], wbd[8];
        hxs::bary(Nd, xg[Nd], wbd);
        hxs::interp_vector(minus ? -1.0 : 1.0, Nd, xg[Nd], wbd, vb);
        for (int b = 0; b < ey; ++b)
            for (int a = 0; a < ex; ++a)
                for (int t = 0; t <= Nd; ++t) {
                    long g;
                    if (d0 == 0) g = id(t, a, b);
                    else if (d0 == 2) g = id(a, b, t);
                    else g = id(a, t, b);
                    for (int c = 0; c < 3; ++c) dS[3 * (a + ex * b) + c] = dS[3 * (a + ex * b) + c] + m->Ja[9 * g + 3 * d0 + c] * vb[t];
                }
        if (minus)
            for (auto &v : dS) v = -v;
        // projection onto the face order (projType 1: first direction, 2: second, 3: both)
        const bool p1 = Ne1 < Nf1, p2 = Ne2 < Nf2;
        if (!p1 && !p2) {
            nrm = dS;
        } else {
            double T1[64], T2[64];
            if (p1) hxs::tset_host(Ne1, Nf1, T1, gauss);
            if (p2) hxs::tset_host(Ne2, Nf2, T2, gauss);
            std::fill(nrm.begin(), nrm.end(), 0.0);
            if (p1 && !p2) {
                for (int j = 0; j < fy; ++j)
                    for (int l = 0; l < ex; ++l)
                        for (int i = 0; i < fx; ++i)
                            for (int c = 0; c < 3; ++c)
                                nrm[3 * (i + fx * j) + c] = nrm[3 * (i + fx * j) + c] + T1[i * ex + l] * dS[3 * (l + ex * j) + c];
            } else if (!p1 && p2) {
                for (int l = 0; l < ey; ++l)
                    for (int j = 0; j < fy; ++j)
                        for (int i = 0; i < fx; ++i)
                            for (int c = 0; c < 3; ++c)
                                nrm[3 * (i + fx * j) + c] = nrm[3 * (i + fx * j) + c] + T2[j * ey + l] * dS[3 * (i + ex * l) + c];
            } else {
                for (int l = 0; l < ey; ++l)
                    for (int j = 0; j < fy; ++j)
                        for (int mm = 0; mm < ex; ++mm)
                            for (int i = 0; i < fx; ++i)
                                for (int c = 0; c < 3; ++c)
                                    nrm[3 * (i + fx * j) + c] =
                                        nrm[3 * (i + fx * j) + c] + T1[i * ex + mm] * T2[j * ey + l] * dS[3 * (mm + ex * l) + c];
            }
        }
        for (int q = 0; q < fx * fy; ++q) {
            const long Q = m->foff[f] + q;
            double *n = &nrm[3 * q];
            const double jac = norm2(n);
            for (int c = 0; c < 3; ++c) n[c] = n[c] / jac;
            m->fJac[Q] = jac;
            for (int c = 0; c < 3; ++c) m->fNormal[3 * Q + c] = n[c];
        }
        // t1 = D x along the first face coordinate, orthogonalised and normalised; t2 = n x t1
        for (int j = 0; j < fy; ++j)
            for (int i = 0; i < fx; ++i) {
                double t[3] = {0.0, 0.0, 0.0};
                for (int l = 0; l < fx; ++l)
                    for (int c = 0; c < 3; ++c) t[c] = t[c] + Dg[Nf1][i * fx + l] * xf[3 * (l + fx * j) + c];
                const long Q = m->foff[f] + i + fx * j;
                const double *n = &m->fNormal[3 * Q];
                const double dp = t[0] * n[0] + t[1] * n[1] + t[2] * n[2];
                for (int c = 0; c < 3; ++c) t[c] = t[c] - dp * n[c];
                const double tn = norm2(t);
                for (int c = 0; c < 3; ++c) t[c] = t[c] / tn;
                for (int c = 0; c < 3; ++c) m->fT1[3 * Q + c] = t[c];
                m->fT2[3 * Q + 0] = n[1] * t[2] - n[2] * t[1];
                m->fT2[3 * Q + 1] = n[2] * t[0] - n[0] * t[2];
                m->fT2[3 * Q + 2] = n[0] * t[1] - n[1] * t[0];
            }
    }
    return HX3D_OK;
}

}  // namespace

int hx3d_front::read_specmesh(hx3d_mesh **out, const char *path, int N, const int *orders, long norders,
                              const ZoneResolver &resolve, bool gauss)
{
    if (!out) return HX3D_EINVAL;
    *out = nullptr;
    hx3d_mesh *m = new (std::nothrow) hx3d_mesh;
    if (!m) return HX3D_ENOMEM;
    *out = m;
    if (!path) return fail(m, HX3D_EINVAL, "no mesh file name");
    try {
        return ::read_specmesh(m, path, N, orders, norders, resolve, gauss);
    } catch (const std::bad_alloc &) {
        return fail(m, HX3D_ENOMEM, "out of host memory reading the mesh");
    } catch (...) {
        return fail(m, HX3D_EINVAL, "mesh file could not be read");
    }
}

extern "C" {

int hx3d_mesh_read_specmesh(const char *path, int N, const int *orders, const char *periodic, hx3d_mesh **out)
{
    return hx3d_mesh_read_specmesh_nodes(path, N, orders, periodic, HX3D_NODES_GAUSSLOBATTO, out);
}

int hx3d_mesh_read_specmesh_nodes(const char *path, int N, const int *orders, const char *periodic, int nodes,
                                  hx3d_mesh **out)
{
    if (nodes != HX3D_NODES_GAUSSLOBATTO && nodes != HX3D_NODES_GAUSS) {
        if (out) *out = nullptr;
        return HX3D_EINVAL;
    }
    // periodic pairs "a=b;c=d": boundary a is periodic, coupled with b; other zones are not periodic
    std::map<std::string, std::string> pairs;
    if (periodic) {
        std::stringstream ss(periodic);
        std::string item;
        auto trim = [](const std::string &s) {
            size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
            return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
        };
        while (std::getline(ss, item, ';')) {
            const size_t eq = item.find('=');
            if (eq != std::string::npos) pairs[lower(trim(item.substr(0, eq)))] = lower(trim(item.substr(eq + 1)));
        }
    }
    auto resolve = [&pairs](const std::string &zone, std::string &c, std::string &) {
        auto it = pairs.find(zone);
        if (it == pairs.end()) return 0;
        c = it->second;
        return 1;
    };
    return hx3d_front::read_specmesh(out, path, N, orders, hx3d_front::ORDERS_UNCHECKED, resolve,
                                     nodes == HX3D_NODES_GAUSS);
}

const char *hx3d_mesh_error(const hx3d_mesh *m) { return m ? m->err.c_str() : "no mesh"; }

long hx3d_mesh_num_nodes(const hx3d_mesh *m) { return m && !m->off.empty() ? m->off.back() : 0; }

long hx3d_mesh_num_face_nodes(const hx3d_mesh *m) { return m && !m->foff.empty() ? m->foff.back() : 0; }

const double *hx3d_mesh_coords(const hx3d_mesh *m) { return m && !m->x.empty() ? m->x.data() : nullptr; }

int hx3d_mesh_fill_desc(const hx3d_mesh *m, hx3d_case_desc *d)
{
    if (!m || !d || m->off.empty()) return HX3D_EINVAL;
    d->K = m->K;
    d->N = m->N.data();
    d->faceID = m->faceID.data();
    d->faceSide = m->faceSide.data();
    d->F = m->F;
    d->fElem = m->fElem.data();
    d->fESide = m->fESide.data();
    d->fRot = m->fRot.data();
    d->Ja = m->Ja.data();
    d->J = m->J.data();
    d->fNormal = m->fNormal.data();
    d->fT1 = m->fT1.data();
    d->fT2 = m->fT2.data();
    d->fJac = m->fJac.data();
    d->nodes = m->nodes;
    return HX3D_OK;
}

int hx3d_mesh_destroy(hx3d_mesh *m)
{
    delete m;
    return HX3D_OK;
}

}  // extern "C"
