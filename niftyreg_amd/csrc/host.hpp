// Host-side helpers shared by the F3D and Aladin drivers (no device use).
#pragma once

#include <vector>
#include "common.hpp"

namespace nrh {

// nifti_mat44_inverse (reg-io/niftilib/nifti1_io.c:1680-1717)
void mat44_inverse(const float *R, float *Q);
// reg_createControlPointGrid<float> (reg-lib/cpu/_reg_localTrans.cpp:21-128)
void create_grid(const nrh_geom& ref, const float spacing[3], nrh_geom& grid, float *identity);

// reg_createSymmetricControlPointGrids<float> without an affine (reg-lib/cpu/_reg_localTrans.cpp:132-358)
void symmetric_grids(const nrh_geom& ref, const nrh_geom& flo, const float spacing[3], nrh_geom& grid, float *fw,
                     float *bw, const float *halfFw = nullptr, const float *halfBw = nullptr);
// operator*(mat44, mat44) of reg-lib/cpu/Maths.hpp:237-246 (double sums, float results)
void mat44_product(const float *A, const float *B, float *R);
// reg_createSymmetricControlPointGrids's half affines; 1 when the grids carry them as extensions
int half_affines(const float A[16], float hf[16], float hb[16]);

struct Level {
    nrh_geom geom;
    std::vector<float> data;
};

// reg_createImagePyramid<float> (reg-lib/cpu/_reg_tools.cpp:1705-1733)
std::vector<Level> pyramid(const nrh_geom& g, const float *data, int levels, int toPerform);

}  // namespace nrh
