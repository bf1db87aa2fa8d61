// FP64 geometric factors: per element J and the affine map's partials, per
// side length and unit normal, and the element-centric arrays the kernels use.
#pragma once
#include <string>
#include <vector>

#include "dg_mesh.h"

namespace dgmi {

// Reference-order geometry (conserv_kernels.cu:338-514, main.cu:84-113):
// geo_e[5K] = J|xr|yr|xs|ys, geo_s[3S] = len|Nx|Ny.  quirks bit 2 rounds
// lengths, normals and inscribed diameters through float as the reference
// does.  Returns the minimum inscribed-circle diameter.
double mesh_geometry(const Mesh& m, int quirks, std::vector<double>& geo_e, std::vector<double>& geo_s);

// Element-centric geometry in the kernel order (element e_new = perm[e_old]'s
// position): per element xr, yr, xs, ys, J; per (element, face) the side's
// nx, ny, len (normal pointing out of the side's left element) and the packed
// neighbour word; vertex coordinates for the initial-condition projection.
struct ElemGeometry {
    int K = 0;
    std::vector<double> xr, yr, xs, ys, J;
    std::vector<double> fnx[3], fny[3], flen[3];
    std::vector<int> nbr[3];
    std::vector<double> V[6];
    double min_r = 0;
    // boundary sides in the order of their index in the neighbour word:
    // element (kernel order), its local face, kind (1 reflecting, 2 outflow,
    // 3 inflow)
    struct BFace {
        int elem, face, kind;
    };
    std::vector<BFace> bfaces;
};

std::string elem_geometry(const Mesh& m, const std::vector<int>& perm, int quirks, ElemGeometry& g);

}  // namespace dgmi
