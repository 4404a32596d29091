// Helpers shared by the reference shims (test infrastructure only): nifti images and
// geometry descriptors for the reference's own classes, built from plain arrays.
#pragma once

#include "_reg_localTrans.h"
#include "_reg_tools.h"
#include <cstring>
#include <string>
#include <vector>

extern "C" {
typedef struct {
    int nx, ny, nz;
    float pixdim[3];  // pixdim[1..3] of the header (0: derive from the matrix columns)
    float sform[16];  // row-major voxel -> mm
} nr_geom;
}

namespace ref_nifti {


inline mat44 to_mat44(const float *m) {
    mat44 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = m[i * 4 + j];
    return r;
}

inline void from_mat44(const mat44& r, float *m) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m[i * 4 + j] = r.m[i][j];
}

// Set sform and qform of an image from a 4x4 voxel->mm matrix.
inline void set_orientation(nifti_image *img, const mat44& m) {
    img->sform_code = 1;
    img->sto_xyz = m;
    img->sto_ijk = nifti_mat44_inverse(m);
    float qb, qc, qd, qx, qy, qz, dx, dy, dz, qfac;
    nifti_mat44_to_quatern(m, &qb, &qc, &qd, &qx, &qy, &qz, &dx, &dy, &dz, &qfac);
    img->qform_code = 1;
    img->quatern_b = qb; img->quatern_c = qc; img->quatern_d = qd;
    img->qoffset_x = qx; img->qoffset_y = qy; img->qoffset_z = qz;
    img->qfac = qfac;
    img->qto_xyz = nifti_quatern_to_mat44(qb, qc, qd, qx, qy, qz, dx, dy, dz, qfac);
    img->qto_ijk = nifti_mat44_inverse(img->qto_xyz);
    img->pixdim[1] = img->dx = dx;
    img->pixdim[2] = img->dy = dy;
    img->pixdim[3] = img->dz = dz;
}

inline NiftiImage make_image(const nr_geom *g, const float *data, int nt = 1, int nu = 1) {
    std::vector<NiftiImage::dim_t> dims{ g->nx, g->ny, g->nz };
    if (nt > 1 || nu > 1) dims.push_back(nt);
    if (nu > 1) dims.push_back(nu);
    NiftiImage img(dims, NIFTI_TYPE_FLOAT32);
    img->pixdim[0] = 1.f;
    set_orientation(img, to_mat44(g->sform));
    if (g->pixdim[0] > 0) {
        img->pixdim[1] = img->dx = g->pixdim[0];
        img->pixdim[2] = img->dy = g->pixdim[1];
        img->pixdim[3] = img->dz = g->pixdim[2];
    }
    if (nt > 1 || nu > 1) { img->pixdim[4] = img->dt = 1.f; }
    if (nu > 1) { img->pixdim[5] = img->du = 1.f; }
    img->scl_slope = 1.f;
    img->scl_inter = 0.f;
    const size_t n = (size_t)g->nx * g->ny * g->nz * nt * nu;
    if (data) std::memcpy(img->data, data, n * sizeof(float));
    else std::memset(img->data, 0, n * sizeof(float));
    return img;
}

// A vector image (field / gradient) laid out like the reference's deformation field.
inline NiftiImage make_vector_image(const nr_geom *g, const float *data) {
    NiftiImage img = make_image(g, data, 1, 3);
    img->intent_code = NIFTI_INTENT_VECTOR;
    img->intent_p1 = DEF_FIELD;
    return img;
}

// A control point grid image with the reference's header conventions.
inline NiftiImage make_grid(const nr_geom *g, const float *data) {
    NiftiImage img = make_vector_image(g, data);
    img->intent_p1 = CUB_SPLINE_GRID;
    img.setIntentName("NREG_TRANS"s);
    return img;
}

inline size_t nvox(const nr_geom *g) { return (size_t)g->nx * g->ny * g->nz; }

// Geometry of an existing image (sform if set, else qform).
inline void geom_of(const nifti_image *img, nr_geom *g) {
    g->nx = img->nx; g->ny = img->ny; g->nz = img->nz;
    g->pixdim[0] = img->dx; g->pixdim[1] = img->dy; g->pixdim[2] = img->dz;
    from_mat44(img->sform_code > 0 ? img->sto_xyz : img->qto_xyz, g->sform);
}

}  // namespace ref_nifti
