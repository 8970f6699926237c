// Host-side interface between the mesh front end (specmesh.cpp) and the .control reader
// (control.cpp).  Not part of the C ABI.
#pragma once

#include <functional>
#include <string>

#include "../../include/hx3d.h"

namespace hx3d_front {

// Boundary condition of one mesh zone (lower-case zone name), as the .control file or the
// "a=b;c=d" pair string gives it: returns 1 and the coupled zone's name when the zone is
// periodic, 0 when it has another boundary condition, or a negative HX3D status with a message.
using ZoneResolver = std::function<int(const std::string &zone, std::string &coupled, std::string &err)>;

// norders: the number of (Nx, Ny, Nz) rows in orders, checked against the mesh, or one of these
constexpr long ORDERS_UNCHECKED = -1;   // one row per element, count not known to the caller
constexpr long ORDERS_ONE_ROW = -2;     // a single row that applies to every element

// Reads a SpecMesh file (ConstructMesh_FromSpecMeshFile_, Read_SpecMesh.f90:36-348) with the
// uniform order N, or the per-element orders when orders is given; the geometry on Gauss nodes
// (gauss) or Gauss-Lobatto nodes.
int read_specmesh(hx3d_mesh **out, const char *path, int N, const int *orders, long norders,
                  const ZoneResolver &resolve, bool gauss = false);

}  // namespace hx3d_front
