/*
 * nr_oracle.h -- CPU restatement of the NiftyReg F3D / Aladin hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: only the tests,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product library
 * (libnrhip.so) never links or calls it.
 *
 * Every function restates the reference CPU algorithm in plain C, following the scalar
 * (non-SSE) summation order of the reference file cited beside it.  Images are planar float
 * arrays, x fastest; vector images hold all x components, then all y, then all z.
 */
#ifndef NR_ORACLE_H
#define NR_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Image geometry: dims, voxel size, voxel->mm (xyz) and mm->voxel (ijk), row-major 4x4. */
typedef struct {
    int nx, ny, nz;
    float dx, dy, dz;
    float xyz[16];
    float ijk[16];
} orc_geom;

/* get_BSplineBasisValues<float> (reg-lib/cpu/_reg_splineBasis.cpp:17-26) */
void orc_bspline_basis(float t, float out[4]);
/* ... with first derivative (reg-lib/cpu/_reg_splineBasis.cpp:29-37) */
void orc_bspline_basis_d1(float t, float val[4], float first[4]);

/* reg_cubic_spline_getDeformationField3D, non-composition (reg-lib/cpu/_reg_localTrans.cpp:1080-1453).
 * The 5-voxel look-up-table path (1110-1304) is used when cpp.d / ref.d == 5 on every axis.
 * mask may be NULL (all active); masked voxels are left untouched on the LUT path and set to 0
 * otherwise, as in the reference. */
void orc_spline_field(const orc_geom *ref, const orc_geom *cpp, const float *cppData,
                      const int *mask, float *field);

/* ResampleImage3D with linear interpolation (reg-lib/cpu/_reg_resampling.cpp:316-464). */
void orc_resample_linear(const orc_geom *flo, const float *floData, const orc_geom *ref,
                         const float *field, const int *mask, float pad, float *warped);

/* TrilinearImageGradient (reg-lib/cpu/_reg_resampling.cpp:1797-1915). */
void orc_image_gradient_linear(const orc_geom *flo, const float *floData, const orc_geom *ref,
                               const float *field, const int *mask, float pad, float *grad);

/* reg_intensityRescale_core<float> for scl_slope 1, scl_inter 0 (reg-lib/cpu/_reg_tools.cpp:93-183). */
void orc_intensity_rescale(float *data, size_t n, float newMin, float newMax);

/* reg_getNmiValue<float>, approximate Parzen window (reg-lib/cpu/_reg_nmi.cpp:160-344).
 * pro / logh: rb*fb + rb + fb doubles; ent: 4 doubles (H_r, H_f, H_rf, total).  Returns NMI. */
double orc_nmi(const float *ref, const float *war, const int *mask, size_t n, int rb, int fb,
               double *pro, double *logh, double *ent);

/* reg_getVoxelBasedNmiGradient3d<float> (reg-lib/cpu/_reg_nmi.cpp:501-593); grad is += . */
void orc_nmi_voxel_gradient(const float *ref, const float *war, const float *warGrad,
                            const int *mask, size_t n, int rb, int fb, const double *logh,
                            const double *ent, double weight, float *grad);

/* reg_tools_kernelConvolution<float> for one volume and one axis (reg-lib/cpu/_reg_tools.cpp:832-1112),
 * kernel types: 2 Gaussian, 3 cubic B-spline.  sigma > 0 is in mm, < 0 in voxels. */
void orc_kernel_convolution_axis(const orc_geom *g, float *data, float sigma, int kernel,
                                 const int *mask, int axis);

/* Compute::ConvolveImage on a 3-component voxel gradient (reg-lib/Compute.cpp:230-264): cubic
 * kernel, x pass with the grid dx, y pass with dy, z pass with dz. */
void orc_convolve_voxel_gradient(const orc_geom *ref, const orc_geom *cpp, float *voxelGrad);

/* reg_voxelCentricToNodeCentric, update = false (reg-lib/cpu/_reg_localTrans.cpp:1540-1684). */
void orc_voxel_to_node(const orc_geom *cpp, const orc_geom *ref, const float *voxelGrad,
                       const float reorient_ijk[16], float weight, float *nodeGrad);

/* reg_spline_approxBendingEnergyValue3D (reg-lib/cpu/_reg_localTrans_regul.cpp:73-153). */
double orc_bending_energy(const orc_geom *cpp, const float *cppData);
/* reg_spline_approxBendingEnergyGradient3D (reg-lib/cpu/_reg_localTrans_regul.cpp:287-451); grad += */
void orc_bending_energy_gradient(const orc_geom *cpp, const float *cppData, float weight, float *grad);
/* set_second_order_bspline_basis_values (reg-lib/cpu/_reg_splineBasis.cpp:292-457) */
void orc_second_order_basis(float xx[27], float yy[27], float zz[27], float xy[27], float yz[27], float xz[27]);

/* reg_getMaximalLength<float> (reg-lib/cpu/_reg_tools.cpp:335-353) */
float orc_max_length(const float *grad, size_t nodes);
/* Compute::NormaliseGradient (reg-lib/Compute.cpp:152-183) */
void orc_normalise_gradient(float *grad, size_t nodes, double maxLength);
/* ConjugateGradient<float>::UpdateGradientValues (reg-lib/Optimiser.cpp:231-317).
 * first != 0 initialises; returns gamma (0 on the first call). */
double orc_cg_update(float *grad, float *g, float *h, size_t n, int first);
/* Compute::UpdateControlPointPosition (reg-lib/Compute.cpp:91-127) */
void orc_update_dof(float *cur, const float *best, const float *grad, float scale, size_t n);

/* reg_affine_getDeformationField3D (reg-lib/cpu/_reg_globalTrans.cpp:74-132) */
void orc_affine_field(const orc_geom *ref, const float affine[16], const int *mask, float *field);

/* mat44 inverse in double, stored in float (nifti_mat44_inverse semantics). */
void orc_mat44_inverse(const float in[16], float out[16]);


/* ---- Aladin ---------------------------------------------------------------------------- */
/* Block grid of 4x4x4 blocks: nb = ceil(n/4) per axis; returns the total block number. */
int orc_block_count(const orc_geom *ref, int nb[3]);
/* Variance of one block, -1 when unusable (reg-lib/cpu/_reg_blockMatching.cpp:70-128). */
float orc_block_variance(const orc_geom *ref, const float *img, const int *mask, int bi, int bj, int bk);
/* _reg_set_active_blocks<float> + initialise_block_matching_method
 * (reg-lib/cpu/_reg_blockMatching.cpp:49-273): totalBlock[b] = rank or -1; returns the active number. */
int orc_active_blocks(const orc_geom *ref, const float *img, const int *mask, int percent, int *totalBlock);
/* block_matching_method3D<float> (reg-lib/cpu/_reg_blockMatching.cpp:455-690); returns the defined count. */
int orc_block_matching(const orc_geom *ref, const float *img, const float *warped, const int *mask,
                       const int *totalBlock, float *refPos, float *warpPos);
/* optimize + optimize_3D (reg-lib/cpu/_reg_blockMatching.cpp:722-821,
 * reg-lib/cpu/_reg_globalTrans.cpp:328-394): LTS update of matrix (row-major, in/out).
 * Returns 0, or -1 when there are too few defined correspondences. */
int orc_lts(const float *refPos, const float *warpPos, int nActive, int percentToKeep, int affine,
            float matrix[16]);


/* ---- LNCC ------------------------------------------------------------------------------ */
/* reg_tools_kernelConvolution<float> on one volume over the three axes (reg-lib/cpu/_reg_tools.cpp:
 * 832-1112, scalar build).  kernel: 2 Gaussian, 3 cubic B-spline; sigma < 0 in voxels. */
void orc_kernel_convolution(const orc_geom *g, float *data, float sigma, int kernel, const int *mask);
/* reg_lncc value, one time point (reg-lib/cpu/_reg_lncc.cpp:195-300); images already in [0, 1]. */
double orc_lncc_value(const orc_geom *g, const float *ref, const float *war, const int *mask, float sigma);
/* reg_getVoxelBasedLnccGradient<float> (reg-lib/cpu/_reg_lncc.cpp:393-512); grad -= */
void orc_lncc_gradient(const orc_geom *g, const float *ref, const float *war, const float *warGrad, const int *mask,
                       float sigma, double weight, float *grad);


/* The displacement round trip reg_spline_approxBendingEnergyGradient3D applies to the control
 * points in place (reg-lib/cpu/_reg_localTrans_regul.cpp:312, 450). */
void orc_displacement_round_trip(const orc_geom *cpp, float *cppData);

/* ---- linear energy ---------------------------------------------------------------------- */
/* set_first_order_basis_values 3-D (reg-lib/cpu/_reg_splineBasis.cpp:182-290) */
void orc_first_order_basis(float bx[27], float by[27], float bz[27]);
/* reg_spline_approxLinearEnergyValue3D (reg-lib/cpu/_reg_localTrans_regul.cpp:559-647) */
double orc_linear_energy(const orc_geom *cpp, const float *cppData);
/* reg_spline_approxLinearEnergyGradient3D (reg-lib/cpu/_reg_localTrans_regul.cpp:739-824); grad += */
void orc_linear_energy_gradient(const orc_geom *cpp, const float *cppData, float weight, float *grad);

/* ---- symmetric Aladin ------------------------------------------------------------------- */
/* nifti_mat44_inverse (reg-io/niftilib/nifti1_io.c:1680-1717), row-major 4x4 */
void orc_nifti_mat44_inverse(const float A[16], float out[16]);
/* NiftyReg::Mat44Avg2 (reg-lib/cpu/Maths.cpp:479-491): expm((logm A + logm B) / 2) */
int orc_mat44_avg2(const float A[16], const float B[16], float out[16]);
/* the matrix update of reg_aladin_sym::UpdateTransformationMatrix (reg-lib/_reg_aladin_sym.cpp:128-157) */
int orc_symmetric_update(float fw[16], float bw[16]);

/* ---- velocity fields (F3D2) --------------------------------------------------------------- */
/* Grids without affine extensions; fields and grids hold mm positions, planar [3][n]. */
/* reg_spline_getFlowFieldFromVelocityGrid (reg-lib/cpu/_reg_localTrans.cpp:3465-3491) */
void orc_vel_flow(const orc_geom *ref, const orc_geom *cpp, const float *vel, float *flow);
/* reg_defField_compose3D, no mask (reg-lib/cpu/_reg_localTrans.cpp:2333-2457); upd in place */
void orc_def_field_compose(const orc_geom *g, const float *def, float *upd);
/* reg_spline_getDefFieldFromVelocityGrid (reg-lib/cpu/_reg_localTrans.cpp:3583-3613); *steps is
 * the grid's intent_p2 (updated when `update` is set or it is 0) */
void orc_vel_def_field(const orc_geom *ref, const orc_geom *cpp, const float *vel, int *steps, int update,
                       float *field);
/* reg_spline_getIntermediateDefFieldFromVelGrid (reg-lib/cpu/_reg_localTrans.cpp:3616-3683);
 * (|steps| + 1) fields */
void orc_vel_intermediate(const orc_geom *ref, const orc_geom *cpp, const float *vel, int steps, float *fields);
/* reg_trilinearResampleGradient (reg-lib/cpu/_reg_resampling.cpp:1547-1766) */
void orc_resample_gradient(const orc_geom *g, const float *grad, const float *def, float pad, float *out);
/* Compute::ExponentiateGradient (reg-lib/Compute.cpp:283-326); grad in place */
void orc_exponentiate_gradient(const orc_geom *ref, const orc_geom *cppBw, const float *velBw, int steps,
                               float *grad);
/* Compute::UpdateVelocityField (reg-lib/Compute.cpp:336-348) */
void orc_update_velocity(float *grid, const float *grad, float scale, size_t nodes, int ox, int oy, int oz);
/* Compute::SymmetriseVelocityFields (reg-lib/Compute.cpp:357-391) */
void orc_symmetrise_velocity(const orc_geom *cpp, float *fw, float *bw);
/* reg_createSymmetricControlPointGrids, no affine (reg-lib/cpu/_reg_localTrans.cpp:132-358) */
void orc_symmetric_grids(const orc_geom *ref, const orc_geom *flo, const float spacing[3], orc_geom *grid,
                         float *fw, float *bw);
/* ---- velocity grids with affine parts: ext0 / ext1 = the grid's extensions 0 and 1, transBw =
 * the backward content's transformation matrix (NULL when absent) */
int orc_half_affines(const float A[16], float hf[16], float hb[16]);
void orc_symmetric_grids_aff(const orc_geom *ref, const orc_geom *flo, const float spacing[3], const float *affine,
                             orc_geom *grid, float *fw, float *bw);
void orc_vel_def_field_aff(const orc_geom *ref, const orc_geom *cpp, const float *vel, const float *ext0,
                           const float *ext1, int *steps, int update, float *field);
void orc_vel_intermediate_aff(const orc_geom *ref, const orc_geom *cpp, const float *vel, const float *ext0,
                              const float *ext1, int steps, float *fields);
void orc_exponentiate_gradient_aff(const orc_geom *ref, const orc_geom *cppBw, const float *velBw,
                                   const float *ext0, const float *ext1, const float *transBw, int steps,
                                   float *grad);
/* reg_voxelCentricToNodeCentric with a grid extension 0 (reg-lib/cpu/_reg_localTrans.cpp:1568-1592) */
void orc_voxel_to_node_ext(const orc_geom *cpp, const orc_geom *ref, const float *voxelGrad,
                           const float reorient_ijk[16], const float ext0[16], float weight, float *nodeGrad);

#ifdef __cplusplus
}
#endif
#endif
