// Triangle mesh in the reference's read_mesh layout plus the element-centric
// Morton-ordered connectivity the HIP kernels use.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dgmi {

// Arrays exactly as the reference's read_mesh returns them
// (src/2d/conservation/conserv.cu:118-192): vertices V1..V3 per element,
// the global side id of each local face (face 0 = V1->V2, 1 = V2->V3,
// 2 = V3->V1), and per side its end points in the left element's
// orientation, left/right element and the local face number on each side.
struct Mesh {
    int K = 0, S = 0;
    std::vector<double> V1x, V1y, V2x, V2y, V3x, V3y;
    std::vector<int> elem_s1, elem_s2, elem_s3;
    std::vector<double> sx1, sy1, sx2, sy2;
    std::vector<int> left_elem, right_elem, left_side, right_side;
    // periodic extent of the domain (0 when the mesh is not periodic); used
    // by the initial conditions to measure distances to the nearest image
    double period_x = 0, period_y = 0;
    double x0 = 0, y0 = 0;
    // true when the element order is already Morton (generated meshes)
    bool ordered = false;
};

// Returns an empty string on success, otherwise the reason.
std::string read_pmsh(const char* path, Mesh& m);

// gmsh 2.2 ASCII mesh (.msh) to the side structure, as the reference's
// converter builds it (src/2d/genmesh.py:11-310): triangles (8-field element
// lines) in file order; sides numbered as first met over (element, local
// face 0..2); the element that first meets a side is its left element;
// 2-node line elements tagged 10000 / 20000 / 30000 mark reflecting /
// outflow / inflow sides (right element -1 / -2 / -3); boundary sides are
// then moved to the front, reflecting first.  With `genmesh_quirks` the
// converter's two latent bugs are reproduced, so shipped .msh/.pmsh pairs
// convert exactly: an element with negative Jacobian swaps the coordinates
// of its first two vertices in the shared vertex list, and the front-sort
// writes the slot of a boundary side's "right element" through a negative
// index (element K + code).  Without it the element's own vertex order is
// swapped, the sort leaves boundary sides' missing right elements alone,
// and a side that belongs to one element but carries no tag is an error.
// With `periodic` (clean mode only) the open untagged sides on the edges of
// the nodes' bounding box are first paired with their translates on the
// opposite edge (the box's width / height becomes the mesh's period), the
// partner side removed and the rest renumbered in order; tagged sides stay
// boundary sides, so a channel periodic in x with walls in y reads too.
std::string read_gmsh(const char* path, Mesh& m, bool genmesh_quirks, bool periodic);
std::string write_pmsh(const char* path, const Mesh& m);

// The advection prototype's 6-field .pmsh (V1..V3 per element) with its sides
// built as the prototype's read_mesh does (src/2d/2dadvec/2dadvec.cu:114-252);
// sides met by one element get right element `open_kind` (-1, -2 or -3).
std::string read_vertex_pmsh(const char* path, Mesh& m, int open_kind);

// n_x * n_y squares on [x0,x1] x [y0,y1], periodic in both directions, each
// split along its (i,j)-(i+1,j+1) diagonal into two counter-clockwise
// triangles.  Lattice vertices off the periodic boundary are moved by
// independent uniform offsets in [-jitter*h, jitter*h] per coordinate, drawn
// from splitmix64(seed) in row-major vertex order (two draws per vertex, x
// then y, for every vertex 0 <= i < n_x, 0 <= j < n_y; draws for boundary
// vertices are made and discarded).  Cells are enumerated in blocks of about
// 8 x 4 cells (ceil(nx/8) x ceil(ny/4) blocks partitioning the grid evenly)
// taken in Morton order of the block index, row-major inside a block;
// the c-th cell owns elements 2c (lower) and 2c+1 (upper).  Sides are
// numbered in order of first appearance over (element, local face), and the
// element that first meets a side is its left element.
std::string make_periodic_jittered(int nx, int ny, double x0, double x1, double y0, double y1, double jitter,
                                   uint64_t seed, Mesh& m);

// Checks that the side/element cross references are consistent.
std::string validate(const Mesh& m);

// Element permutation by the Morton code of the centroid (new -> old).
std::vector<int> morton_order(const Mesh& m);

uint64_t splitmix64(uint64_t& state);

}  // namespace dgmi
