/*
 * niftyreg_hip.h -- C ABI of libnrhip.so, the MI355X (gfx950) implementation of NiftyReg's
 * F3D objective/gradient step and Aladin's block-matching + LTS step.
 *
 * Plain C types only: images are planar float arrays (x fastest; a vector image holds all x
 * components, then all y, then all z, as the reference's host nifti layout), geometries are
 * nrh_geom values, handles are opaque.  Every call returns NRH_OK (0) or an error code; the
 * message of the last error of the calling thread is returned by nrh_last_error().
 * Each declaration names the reference interface it replaces (paths relative to the
 * NiftyReg source tree).  A handle owns one HIP stream; calls on different handles may run
 * concurrently from different host threads, calls on one handle must be serialised.
 */
#ifndef NIFTYREG_HIP_H
#define NIFTYREG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRH_OK 0
#define NRH_ERR_ARG 1          /* invalid argument (NR_FATAL_ERROR in the reference) */
#define NRH_ERR_HIP 2          /* HIP runtime error */
#define NRH_ERR_STATE 3        /* call out of order (e.g. gradient before histogram) */
#define NRH_ERR_NOMEM 4        /* device allocation failed */
#define NRH_ERR_UNSUPPORTED 5  /* configuration the library does not implement */
#define NRH_ERR_NUMERIC 6      /* degenerate data (no active voxel / block, singular fit) */

/* Geometry of a 3-D image as the reference reads it from a nifti header.
 * xyz / ijk: the matrix pair the reference selects (sform if sform_code > 0, else qform;
 * reg-lib/Content.h:31-36), row-major, voxel->mm and mm->voxel.  dx, dy, dz: pixdim[1..3].
 * sform: 1 when xyz/ijk are the sform (changes the node-spacing ratio of
 * reg_voxelCentricToNodeCentric, reg-lib/cpu/_reg_localTrans.cpp:1595-1604). */
typedef struct nrh_geom {
    int nx, ny, nz;
    float dx, dy, dz;
    float xyz[16];
    float ijk[16];
    int sform;
} nrh_geom;

/* ------------------------------------------------------------------ library */
const char *nrh_last_error(void);
int nrh_version(void);
/* Device selection; replaces CudaContext::SetCudaIdx (reg-lib/cuda/CudaContext.cpp:17-25) and
 * Platform::SetGpuIdx (reg-lib/Platform.cpp:66-93). */
int nrh_device_count(int *count);
int nrh_init(int device);

/* ------------------------------------------------------------------ host helpers (no device use) */
/* nifti_mat44_inverse (reg-io/niftilib/nifti1_io.c:1680-1717) */
int nrh_mat44_inverse(const float in[16], float out[16]);
/* reg_createControlPointGrid<float> (reg-lib/cpu/_reg_localTrans.cpp:21-128): grid geometry and
 * the identity control-point positions (planar, 3*G floats; may be NULL to query the size). */
int nrh_create_control_point_grid(const nrh_geom *ref, const float spacing_mm[3], nrh_geom *grid,
                                  float *identity);
/* reg_spline_refineControlPointGrid (reg-lib/cpu/_reg_localTrans.cpp:1809-2190) with the level's
 * reference image: half the spacing, cubic B-spline subdivision of the positions.  Call with
 * out_data NULL to obtain the refined geometry first. */
int nrh_refine_control_point_grid(const nrh_geom *ref, const nrh_geom *grid, const float *data, nrh_geom *out,
                                  float *out_data);
/* reg_downsampleImage<float>(image, 1, axes) (reg-lib/cpu/_reg_tools.cpp:1343-1467): Gaussian
 * smoothing (0.7355 voxel) and nearest-neighbour decimation by 2 along the axes flagged in
 * axes[0..2].  Call with out_data NULL to obtain the output geometry first. */
int nrh_downsample_image(const nrh_geom *in, const float *data, const int axes[3], nrh_geom *out, float *out_data);
/* reg_createSymmetricControlPointGrids<float> (reg-lib/cpu/_reg_localTrans.cpp:132-358): the
 * forward and backward velocity grids of reg_f3d2, axis-aligned (sform only) over the union of
 * both images' corners, each image first carried half-way by its half affine when `affine`
 * (reference -> floating, row-major 4x4; NULL: none) is given.  fw / bw receive the identity
 * positions (3*G floats each; NULL to query the geometry).  The half affines, which the
 * reference stores in the grids' extensions, come from nrh_symmetric_half_affines. */
int nrh_create_symmetric_control_point_grids(const nrh_geom *ref, const nrh_geom *flo, const float spacing_mm[3],
                                             const float *affine, nrh_geom *grid, float *fw, float *bw);
/* The half affines of reg_createSymmetricControlPointGrids (reg-lib/cpu/_reg_localTrans.cpp:
 * 150-166, NiftyReg::Mat44Logm / Mat44Expm of reg-lib/cpu/Maths.cpp:414-478): half_fw =
 * expm(logm(affine) / 2), half_bw = expm(logm(affine^-1) / 2), rounded through float as the
 * reference.  *extensions = 1 when the reference stores them as the grids' extensions 0 and 1
 * (both halves differ from the identity, :320-353), else 0. */
int nrh_symmetric_half_affines(const float affine[16], float half_fw[16], float half_bw[16], int *extensions);

/* ------------------------------------------------------------------ F3D */
typedef struct nrh_f3d nrh_f3d;

typedef struct nrh_f3d_params {
    int ref_bins;             /* internal NMI bins = user bins + 4 (reg-lib/_reg_base.cpp:567-585) */
    int flo_bins;
    float padding;            /* warped padding value (NaN by default, reg-lib/_reg_base.cpp:46) */
    float similarity_weight;  /* reg_base::similarityWeight after CheckParameters */
    float be_weight;          /* bending-energy weight of the current level */
    float le_weight;          /* linear-energy weight of the current level */
    int image_storage;        /* 0: float32 images in HBM, 1: float16 storage, float32 compute */
    int measure;              /* NRH_MEASURE_NMI or NRH_MEASURE_LNCC (reg-lib/MeasureCreator.hpp:8) */
    float lncc_sigma;         /* LNCC Gaussian standard deviation, < 0: voxels (reg_lncc default -5) */
    int field_order;          /* NRH_FIELD_ORDER_*: which reference build's float summation orders the
                                 device reproduces -- the B-spline field wherever it decides a floor
                                 (reg-lib/cpu/_reg_localTrans.cpp:1110-1451), the velocity grid's
                                 flow (:903-1079) and the density-normalised convolution
                                 (reg-lib/cpu/_reg_tools.cpp:1018-1060) */
} nrh_f3d_params;

#define NRH_MEASURE_NMI 0
#define NRH_MEASURE_LNCC 1

/* The reference sums the 64 weighted control points of a voxel in float, in an order that
 * depends on its build: sequentially (USE_SSE off), or in four x lanes whose partial sums are
 * added pairwise (the 5-voxel look-up-table path with SSE3) or left to right (the general path).
 * Positions that land within rounding of a voxel plane of the floating image decide the
 * trilinear cell, and at the image border whether the voxel is inside, by that last bit. */
#define NRH_FIELD_ORDER_SCALAR 0
#define NRH_FIELD_ORDER_SSE 1

/* F3dContent + DefContent + Content construction (reg-lib/F3dContent.cpp:4-19,
 * reg-lib/DefContent.cpp:5-15, reg-lib/Content.cpp:5-22) and the device upload of
 * CudaF3dContent (reg-lib/cuda/CudaF3dContent.cpp:4-30).  mask: int per reference voxel,
 * > -1 active (NULL = all active).  cpp: grid geometry; the positions are set with
 * nrh_f3d_set_control_points. */
int nrh_f3d_create(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                   const int *mask, const nrh_geom *cpp, const nrh_f3d_params *params, nrh_f3d **out);
int nrh_f3d_destroy(nrh_f3d *h);
int nrh_f3d_set_params(nrh_f3d *h, const nrh_f3d_params *params);
/* F3dContent::UpdateControlPointGrid / GetControlPointGrid (reg-lib/F3dContent.h:17-25) */
int nrh_f3d_set_control_points(nrh_f3d *h, const float *planar);
int nrh_f3d_get_control_points(nrh_f3d *h, float *planar);
/* reg_nmi::InitialiseMeasure's reg_intensityRescale of both images (reg-lib/cpu/_reg_nmi.cpp:91-158,
 * reg-lib/cpu/_reg_tools.cpp:93-208): rescale reference and floating to [lo, hi] on the device. */
int nrh_f3d_rescale_intensities(nrh_f3d *h, float ref_lo, float ref_hi, float flo_lo, float flo_hi);
int nrh_f3d_get_images(nrh_f3d *h, float *ref, float *flo);

/* Per-operation entry points, one per Compute / reg_nmi virtual (reg-lib/Compute.h:11-44,
 * reg-lib/cpu/_reg_measure.h:59-86).  They read and write device-resident buffers of the
 * handle; the nrh_f3d_download_* calls copy those buffers to the host. */
int nrh_f3d_deformation_field(nrh_f3d *h);                 /* Compute::GetDeformationField(false, true) */
int nrh_f3d_resample(nrh_f3d *h);                          /* Compute::ResampleImage(1, padding) */
int nrh_f3d_image_gradient(nrh_f3d *h);                    /* Compute::GetImageGradient(1, padding, 0) */
int nrh_f3d_nmi_value(nrh_f3d *h, double *nmi, double entropies[4]); /* reg_nmi::GetSimilarityMeasureValueFw */
int nrh_f3d_zero_voxel_gradient(nrh_f3d *h);               /* DefContent::ZeroVoxelBasedMeasureGradient */
int nrh_f3d_nmi_voxel_gradient(nrh_f3d *h, double weight); /* reg_nmi::GetVoxelBasedSimilarityMeasureGradientFw */
/* reg_lncc (reg-lib/cpu/_reg_lncc.cpp): GetSimilarityMeasureValueFw (:303-345, the local
 * statistics of UpdateLocalStatImages :195-245 and the mean |local correlation| of
 * reg_getLnccValue :247-300) on the handle's warped image, and
 * GetVoxelBasedSimilarityMeasureGradientFw (:514-577, reg_getVoxelBasedLnccGradient :393-512):
 * voxel gradient -= ... with the given time-point weight.  The images must already be rescaled
 * to [0, 1] (reg_lncc::InitialiseMeasure, :84-105). */
int nrh_f3d_lncc_value(nrh_f3d *h, double *value);
int nrh_f3d_lncc_voxel_gradient(nrh_f3d *h, double weight);
int nrh_f3d_convolve_voxel_gradient(nrh_f3d *h);           /* Compute::ConvolveImage (private) */
int nrh_f3d_voxel_to_node(nrh_f3d *h, float weight);       /* Compute::VoxelCentricToNodeCentric */
int nrh_f3d_zero_node_gradient(nrh_f3d *h);                /* F3dContent::ZeroTransformationGradient */
/* Compute::ApproxLinearEnergy / ApproxLinearEnergyGradient (reg-lib/Compute.cpp:52-62):
 * reg_spline_approxLinearEnergyValue3D / Gradient3D (reg-lib/cpu/_reg_localTrans_regul.cpp:559-824). */
int nrh_f3d_linear_energy(nrh_f3d *h, double *value);
int nrh_f3d_linear_energy_gradient(nrh_f3d *h, float weight);
int nrh_f3d_bending_energy(nrh_f3d *h, double *value);     /* Compute::ApproxBendingEnergy */
int nrh_f3d_bending_energy_gradient(nrh_f3d *h, float weight); /* Compute::ApproxBendingEnergyGradient */
int nrh_f3d_max_length(nrh_f3d *h, double *value);         /* Compute::GetMaximalLength */
int nrh_f3d_normalise_gradient(nrh_f3d *h, double max_length); /* Compute::NormaliseGradient */

int nrh_f3d_download_field(nrh_f3d *h, float *planar);          /* 3*N */
int nrh_f3d_download_warped(nrh_f3d *h, float *warped);         /* N */
int nrh_f3d_download_warped_gradient(nrh_f3d *h, float *planar);/* 3*N */
int nrh_f3d_download_voxel_gradient(nrh_f3d *h, float *planar); /* 3*N */
int nrh_f3d_download_node_gradient(nrh_f3d *h, float *planar);  /* 3*G */
int nrh_f3d_upload_node_gradient(nrh_f3d *h, const float *planar);
/* reg_nmi's jointHistogramPro / jointHistogramLog (rb*fb + rb + fb doubles each) */
int nrh_f3d_download_histogram(nrh_f3d *h, double *pro, double *logh);

/* ---- velocity grids (reg_f3d2) ----
 * A handle whose control points form a cubic B-spline velocity grid (reg_f3d2's forward grid on
 * the reference -> floating handle, its backward grid on a second handle created with the two
 * images swapped), with its affine parts set by nrh_f3d_set_velocity_affines. */
/* The grid set-up of reg_f3d2::Initialise (reg-lib/_reg_f3d2.cpp:585-590): intent_p1 =
 * SPLINE_VEL_GRID, intent_p2 = steps (squarings of the exponentiation; 0 lets the next
 * nrh_f3d_def_field_from_velocity choose). */
int nrh_f3d_set_velocity_grid(nrh_f3d *h, int steps);
/* the grid's intent_p2 (read after an exponentiation that updated it; set to copy the forward
 * step count onto the backward grid, reg-lib/_reg_f3d2.cpp:146-152) */
int nrh_f3d_get_velocity_steps(nrh_f3d *h, int *steps);
int nrh_f3d_set_velocity_steps(nrh_f3d *h, int steps);
/* The affine parts of a velocity grid (row-major 4x4 each, NULL when absent): ext0 / ext1 are the
 * grid's NIfTI extensions 0 and 1 (reg_f3d2 stores the half affine in both).  ext0 is applied
 * before the spline (reg-lib/cpu/_reg_localTrans.cpp:1478-1487), removed before the scaling and
 * squaring and restored after it (:3508-3515, 3569-3573), and maps the grid into the reference
 * space of the node gradient (:1568-1592); ext1 is composed last (:3576-3577).  transformation
 * is the content's transformation matrix (Content::GetTransformationMatrix: reg_f3d2 gives the
 * forward content the input affine and the backward one its inverse), whose displacement the
 * opposite direction's gradient exponentiation removes from each intermediate field
 * (reg-lib/Compute.cpp:303-316). */
int nrh_f3d_set_velocity_affines(nrh_f3d *h, const float *ext0, const float *ext1, const float *transformation);
/* Compute::GetDefFieldFromVelocityGrid (reg-lib/Compute.cpp:223-228): scaling and squaring of
 * the grid into the handle's deformation field (then nrh_f3d_resample / _image_gradient use it). */
int nrh_f3d_def_field_from_velocity(nrh_f3d *h, int update_step_number);
/* Compute::ExponentiateGradient (reg-lib/Compute.cpp:283-326): h's voxel-based gradient carried
 * along the intermediate fields of bw's velocity grid and averaged. */
int nrh_f3d_exponentiate_gradient(nrh_f3d *h, nrh_f3d *bw);
/* Compute::UpdateVelocityField (reg-lib/Compute.cpp:336-348): grid += node gradient * scale. */
int nrh_f3d_update_velocity_field(nrh_f3d *h, float scale, int optimise_x, int optimise_y, int optimise_z);
/* Compute::SymmetriseVelocityFields (reg-lib/Compute.cpp:357-391) of the two grids. */
int nrh_f3d_symmetrise_velocity_fields(nrh_f3d *fw, nrh_f3d *bw);
/* Compute::BchUpdate (reg-lib/Compute.cpp:350-355): the reference's compute_BCH_update stops
 * with "needs updating" (reg-lib/cpu/_reg_localTrans.cpp:3828); this returns NRH_ERR_UNSUPPORTED. */
int nrh_f3d_bch_update(nrh_f3d *h, float scale, int bch_update_value);
/* Compute::DefFieldCompose (reg-lib/Compute.cpp:393-396): the handle's deformation field is
 * replaced by `field` (host, planar 3*voxels(geom), positions) evaluated at its positions. */
int nrh_f3d_compose_field(nrh_f3d *h, const nrh_geom *geom, const float *field);
/* Compute::ResampleGradient (reg-lib/Compute.cpp:399-405), linear only: the voxel-based gradient
 * resampled through the deformation field, into out (host, 3*N). */
int nrh_f3d_resample_gradient(nrh_f3d *h, int interpolation, float padding, float *out);
/* Deformation / voxel-gradient upload for the per-operation checks (DefContent::SetDeformationField,
 * F3dContent::GetVoxelBasedMeasureGradient written by the caller). */
int nrh_f3d_upload_field(nrh_f3d *h, const float *planar);
int nrh_f3d_upload_voxel_gradient(nrh_f3d *h, const float *planar);
/* reg_f3d2::GetObjectiveFunctionValue (reg-lib/_reg_f3d2.cpp:540-560) over the forward handle fw
 * (reference -> floating) and the backward handle bw (floating -> reference, bin counts
 * swapped), NMI: both exponentiations (the forward step number updated, the backward one
 * copied), both warps, symmetric NMI, bending / linear energy of both grids.  terms: wMeasure,
 * wBE, wLE (may be NULL).  The two directions run on the two handles' streams concurrently. */
int nrh_f3d2_objective(nrh_f3d *fw, nrh_f3d *bw, double *value, double terms[3]);
/* reg_f3d2::GetObjectiveFunctionGradient (reg-lib/_reg_f3d2.cpp:417-440): voxel-based NMI
 * gradients of both directions, each exponentiated along the other's grid, convolved and
 * projected; penalty gradients of both grids.  With max_length non-NULL, reg_f3d2::
 * NormaliseGradient (:373-395): both divided by the larger maximal length, which is returned. */
int nrh_f3d2_gradient(nrh_f3d *fw, nrh_f3d *bw, double *max_length);
/* ConjugateGradient<T>::UpdateGradientValues with the backward arrays of reg_f3d2
 * (reg-lib/Optimiser.cpp:231-310): both directions' gradients in one conjugate-gradient step. */
int nrh_f3d2_conjugate_gradient(nrh_f3d *fw, nrh_f3d *bw, int first);

/* Fused objective: reg_f3d::GetObjectiveFunctionValue (reg-lib/_reg_f3d.cpp:443-463) in one
 * device sweep (field, warp and joint histogram fused; no field or warped image is written).
 * terms: wMeasure, wBE, wLE (may be NULL). */
int nrh_f3d_objective(nrh_f3d *h, double *value, double terms[3]);
/* Fused gradient: reg_f3d::GetObjectiveFunctionGradient (reg-lib/_reg_f3d.cpp:553-577) with the
 * similarity gradient, convolution and node projection in one device sweep.  The histogram
 * is recomputed unless the last nrh_f3d_objective ran on the same control points.  When
 * max_length is non-NULL the gradient is then normalised as reg_f3d::NormaliseGradient
 * (reg-lib/_reg_f3d.cpp:395-411) and the maximal length returned; with max_length NULL the call
 * only enqueues the work on the handle's stream (the downloads wait for it). */
int nrh_f3d_gradient(nrh_f3d *h, double *max_length);
/* The similarity stages of the two calls above, for callers that assemble the objective
 * themselves (the HIP platform plugin):
 * nrh_f3d_similarity_value: reg_base::WarpFloatingImage + the measure value of
 * ComputeSimilarityMeasure (reg-lib/_reg_base.cpp:469-495, :688-709) at the current control
 * points, unweighted (NMI, or LNCC); NMI runs the fused histogram sweep when the geometry allows it
 * (no field or warped image written).
 * nrh_f3d_similarity_gradient: reg_f3d::GetSimilarityMeasureGradient (reg-lib/_reg_f3d.cpp:
 * 352-360): the voxel-based measure gradient (time-point weight 1), convolved and extracted at the
 * nodes with `weight` into the node gradient (overwritten); NMI runs the fused gradient sweep and
 * the separable projection when the geometry allows it.  Only enqueues. */
int nrh_f3d_similarity_value(nrh_f3d *h, double *value);
int nrh_f3d_similarity_gradient(nrh_f3d *h, float weight);
/* 1 when the handle's geometry takes the fused sweeps (integer grid spacing >= 4 voxels with
 * nodes on voxels), 0 when objective / gradient run the per-operation chain. */
int nrh_f3d_fused_available(nrh_f3d *h, int *available);
/* The fused objective sweep with its deformation field also written to the handle's field
 * buffer (read with nrh_f3d_download_field): checks the separable field evaluation against
 * reg_spline_getDeformationField (reg-lib/cpu/_reg_localTrans.cpp:1457-1538). */
int nrh_f3d_fused_field(nrh_f3d *h);
/* HIP-event timing of the two fused sweeps on the handle's stream (no reference counterpart;
 * the reference times whole calls with steady_clock, reg-apps/reg_benchmark.cpp:139-163).
 * enable = k > 0 times every k-th sweep of each kind (1: every sweep; the events around a timed
 * sweep cost a few microseconds of stream time), 0 stops timing.  Enabling resets the totals;
 * nrh_f3d_sweep_times returns the accumulated milliseconds and launch counts of the timed
 * [histogram sweeps, gradient sweeps]. */
int nrh_f3d_set_profiling(nrh_f3d *h, int enable);
int nrh_f3d_sweep_times(nrh_f3d *h, double ms[2], long counts[2]);

/* Device-side optimiser state (replaces CudaOptimiser / CudaConjugateGradient,
 * reg-lib/cuda/CudaOptimiser.cu:72-86,198-304): best DOF copy, PR conjugate gradient update
 * (first != 0 initialises), DOF update cur = best + scale * grad. */
int nrh_f3d_store_best(nrh_f3d *h);
int nrh_f3d_restore_best(nrh_f3d *h);
int nrh_f3d_conjugate_gradient(nrh_f3d *h, int first);
int nrh_f3d_update_control_points(nrh_f3d *h, float scale);

/* Full F3D registration on one pair: reg_base<float>::Run (reg-lib/_reg_base.cpp:727-819) with
 * reg_f3d's level set-up (reg-lib/_reg_f3d.cpp:92-133,142-290), NMI, bending energy, CG and
 * line search.  The pyramid and the control-point grid are built on the host; all
 * evaluations run on the device.  cpp_out receives the final grid (3*G floats, G from
 * grid_out), final_nmi the NMI of the last level. */
typedef struct nrh_f3d_config {
    int bins;              /* user bins (64) */
    float be_weight;       /* -be (0.001) */
    float le_weight;       /* -le (the reference's default is 0.01) */
    float spacing;         /* -sx: negative = voxels (-5) */
    int levels;            /* -ln (3) */
    int levels_to_perform; /* -lp (0 = all) */
    int max_iterations;    /* -maxit: final-level budget (150) */
    float padding;         /* -pad (NaN) */
    int use_pyramid;       /* 1 unless -nopy */
    int image_storage;     /* 0 float32, 1 float16 */
    int measure;           /* NRH_MEASURE_NMI (--nmi) or NRH_MEASURE_LNCC (--lncc) */
    float lncc_sigma;      /* --lncc value: Gaussian standard deviation, < 0 voxels (-5) */
    const float *affine;   /* -aff: reference -> floating 4x4, row-major (NULL: none).  reg_f3d
                              starts its grid at the affine's positions (reg-lib/_reg_f3d.cpp:
                              190-191); reg_f3d2 splits it into half affines on both grids */
    int field_order;       /* NRH_FIELD_ORDER_*: the reference build whose field rounding is reproduced */
} nrh_f3d_config;
int nrh_f3d_register(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                     const nrh_f3d_config *cfg, nrh_geom *grid_out, float *cpp_out, size_t cpp_capacity,
                     double *final_nmi, int *iterations);
/* A whole symmetric registration: reg_f3d2's run (reg_base<T>::Run with the reg_f3d2 overrides,
 * reg-lib/_reg_f3d2.cpp, reg-lib/_reg_base.cpp:727-819) with NMI or LNCC and an optional affine:
 * symmetric velocity grids at the coarsest level (spacing < 0: voxels of the mean voxel size),
 * refined by subdivision per level; every evaluation on the device.  fw_out / bw_out (capacity
 * floats each, may be NULL) receive the final forward / backward velocity grids, final_nmi the
 * sum of both directions' NMI. */
int nrh_f3d2_register(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                      const nrh_f3d_config *cfg, nrh_geom *grid_out, float *fw_out, float *bw_out, size_t capacity,
                      double *final_nmi, int *iterations);

/* ------------------------------------------------------------------ batches of F3D pairs */
/* No counterpart in the reference, which runs one registration per process on a process-wide
 * device context (reg-lib/_reg_base.cpp:727-819, reg-lib/cuda/CudaContext.hpp:9-40).  A batch
 * evaluates many reg_f3d handles of different image and grid sizes with the same launches:
 * persistent objective / gradient sweeps over a (pair, tile) work list and one launch per small
 * kernel for all pairs, on the batch's own stream.  Every pair's arithmetic and reduction order
 * are those of the pair alone, so each pair's objective, node gradient and max length are
 * bit-identical to nrh_f3d_objective / nrh_f3d_gradient on that handle.  Pairs must take the
 * fused path (integer grid spacing >= 4 voxels, nodes on voxels), use NMI with a bending-energy
 * term, no linear energy and no mask, and share one image storage type. */
typedef struct nrh_batch nrh_batch;
/* The handles stay owned by the caller and must outlive the batch; while it exists their work
 * runs on the batch's stream, and calls on them must not overlap calls on the batch. */
int nrh_batch_create(nrh_f3d *const *pairs, int n, nrh_batch **out);
int nrh_batch_destroy(nrh_batch *b);
/* reg_f3d::GetObjectiveFunctionValue of every pair (reg-lib/_reg_f3d.cpp:443-463): values[n] and
 * the NMI nmi[n] (either may be NULL). */
int nrh_batch_objective(nrh_batch *b, double *values, double *nmi);
/* reg_f3d::GetObjectiveFunctionGradient of every pair (reg-lib/_reg_f3d.cpp:553-577), normalised
 * by each pair's max length (NormaliseGradient, reg-lib/_reg_f3d.cpp:395-411) when max_lengths is
 * given, which then receives them. */
int nrh_batch_gradient(nrh_batch *b, double *max_lengths);
/* The persistent sweeps' grid: workgroups per XCD (8 XCDs). */
int nrh_batch_set_sweep_blocks(nrh_batch *b, int blocks_per_xcd);
/* Times the histogram sweep for 32, 64, 128 and 256 workgroups per XCD (ms[4], per sweep) and
 * keeps the fastest, returned in *blocks_per_xcd. */
int nrh_batch_autotune(nrh_batch *b, int *blocks_per_xcd, double ms[4]);

typedef struct nrh_batch_stats {
    double seconds;          /* wall time of the whole batched registration */
    double device_seconds;   /* of which: replays of the captured iteration graphs */
    int graph_replays;       /* iteration graph launches over all levels */
    int blocks_per_xcd[8];   /* autotuned persistent grid of each level (workgroups per XCD) */
    int tile[3];             /* voxel tile of the sweeps (x, y, z) */
} nrh_batch_stats;

/* reg_f3d's whole registration (as nrh_f3d_register, same configuration for every pair) for n
 * pairs at once: each pair takes exactly the decisions it takes alone, so its final grid, NMI and
 * iteration count equal nrh_f3d_register's on the same pair; all pairs' evaluations share the
 * batch launches, and each level's iterations replay one captured HIP graph.  grids_out[n];
 * cpp_out[q] (capacity[q] floats, may be NULL), final_nmi[n] and iterations[n] may be NULL;
 * stats may be NULL.  NMI with a bending-energy term, no linear energy. */
int nrh_f3d_register_batch(int n, const nrh_geom *refs, const float *const *ref_data, const nrh_geom *flos,
                           const float *const *flo_data, const nrh_f3d_config *cfg, nrh_geom *grids_out,
                           float *const *cpp_out, const size_t *capacity, double *final_nmi, int *iterations,
                           nrh_batch_stats *stats);

/* ------------------------------------------------------------------ Aladin (affine) */
/* One pyramid level of reg_aladin on the device: the HIP counterpart of AladinContent
 * (reg-lib/AladinContent.cpp:6-29) with its block-matching parameters
 * (_reg_blockMatchingParam, reg-lib/cpu/_reg_blockMatching.h:32-62).  The current affine
 * matrix (floating mm -> reference mm direction of reg_aladin's affineTransformation) lives on
 * the device, so warp -> block matching -> LTS iterations run without host round trips. */
typedef struct nrh_aladin nrh_aladin;

typedef struct nrh_aladin_params {
    int block_percent;  /* percentage of blocks used (reg_aladin -%v, 50) */
    int inlier_lts;     /* percentage of inliers kept by the LTS (reg_aladin -%i, 50) */
    float padding;      /* warped padding value (NaN) */
} nrh_aladin_params;

/* AladinContent construction: uploads both images (and the optional reference mask, int per
 * voxel, < 0 excluded), then selects the active blocks as initialise_block_matching_method /
 * _reg_set_active_blocks<float> (reg-lib/cpu/_reg_blockMatching.cpp:213-273,49-206): block
 * variances on the device, heap-sort ranking (reg-lib/cpu/Maths.cpp:114-153) on the host.
 * Fails with NRH_ERR_NUMERIC when fewer than 2 blocks are usable. */
int nrh_aladin_create(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                      const int *mask, const nrh_aladin_params *params, nrh_aladin **out);
int nrh_aladin_destroy(nrh_aladin *h);
/* counts = {totalBlockNumber, activeBlockNumber, definedActiveBlockNumber of the last block
 * matching}; total_block (may be NULL) receives _reg_blockMatchingParam::totalBlock: the rank of
 * each block among the active ones, or -1. */
int nrh_aladin_blocks(nrh_aladin *h, int counts[3], int *total_block);
/* The current affine (row-major 4x4; reg_aladin::affineTransformation). */
int nrh_aladin_set_matrix(nrh_aladin *h, const float m[16]);
int nrh_aladin_get_matrix(nrh_aladin *h, float m[16]);
/* reg_aladin::GetWarpedImage (reg-lib/_reg_aladin.cpp:256-260): AffineDeformationFieldKernel
 * (reg_affine_getDeformationField, reg-lib/cpu/_reg_globalTrans.cpp:74-175) fused with the
 * trilinear ResampleImageKernel (reg-lib/cpu/_reg_resampling.cpp:316-464); the field is not
 * stored. */
int nrh_aladin_warp(nrh_aladin *h);
int nrh_aladin_upload_warped(nrh_aladin *h, const float *warped);
int nrh_aladin_download_warped(nrh_aladin *h, float *warped);
/* BlockMatchingKernel: block_matching_method3D<float> (reg-lib/cpu/_reg_blockMatching.cpp:455-690),
 * 4x4x4 blocks, capture range 3 voxels, step 1. */
int nrh_aladin_block_matching(nrh_aladin *h);
/* _reg_blockMatchingParam::referencePosition / warpedPosition (3 floats per active block, in
 * rank order, mm; NaN warped position for a block without a match). */
int nrh_aladin_download_positions(nrh_aladin *h, float *ref_pos, float *warp_pos, int *defined);
/* LtsKernel: optimize (reg-lib/cpu/_reg_blockMatching.cpp:722-821) -> optimize_3D
 * (reg-lib/cpu/_reg_globalTrans.cpp:328-394), affine (EstimateAffineLeastSquares) or rigid
 * (EstimateRigidLeastSquares, reg-lib/cpu/Maths.cpp:325-390), in FP64 in one device workgroup;
 * updates the handle's matrix.  Fails with NRH_ERR_NUMERIC when fewer than 8 (affine) / 4
 * (rigid) correspondences are defined, leaving the matrix unchanged. */
int nrh_aladin_lts(nrh_aladin *h, int affine);
/* The same LTS on host arrays (n correspondences, NaN warped x = undefined); m is in/out. */
int nrh_lts(const float *ref_pos, const float *warp_pos, int n, int inlier_lts, int affine, float m[16]);
/* No counterpart in the reference.  nrh_lts for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): ref_pos and
 * warp_pos are device arrays of 3 n floats, matrix a device array of 16 floats holding the current
 * matrix and receiving the fit, status a device int receiving 0 (fit), 1 (too few
 * correspondences) or 2 (singular fit; the matrix is left unchanged).  n <= 262144.  The call
 * only enqueues work on `stream` (a hipStream_t, NULL for the null stream) and allocates nothing:
 * its scratch is a 7.6 MB static device array of the library, one per GPU, kept until the process
 * ends, so calls that overlap on one GPU race on it. */
int nrh_lts_static(const float *ref_pos, const float *warp_pos, int n, int inlier_lts, int affine, float *matrix,
                   int *status, void *stream);

/* The matrix step of reg_aladin_sym::UpdateTransformationMatrix (reg-lib/_reg_aladin_sym.cpp:
 * 140-153), run after both handles' LTS fits: forward <- Mat44Avg2(forward, inverse(backward)),
 * backward <- Mat44Avg2(inverse(forward), backward) (reg-lib/cpu/Maths.cpp:479-491, the
 * log-Euclidean mean), bottom rows reset; on the device, in FP64.  backward is the handle whose
 * reference is forward's floating image.  NRH_ERR_NUMERIC if a logarithm does not exist. */
int nrh_aladin_symmetric_update(nrh_aladin *forward, nrh_aladin *backward);

/* Full affine registration: reg_aladin<float>::Run (reg-lib/_reg_aladin.cpp:315-364) with
 * InitialiseRegistration's pyramid and centre alignment (:88-245), or reg_aladin_sym<float>
 * (reg-lib/_reg_aladin_sym.cpp) when symmetric is set. */
typedef struct nrh_aladin_config {
    int levels;             /* -ln (3) */
    int levels_to_perform;  /* -lp (0 = all) */
    int max_iterations;     /* -maxit (5; doubled at the first level) */
    int block_percent;      /* -%v (50) */
    int inlier_lts;         /* -%i (50) */
    int rigid;              /* rigid stage (1 unless -affDirect) */
    int affine;             /* affine stage (1 unless -rigOnly) */
    int align_centre;       /* 1 unless -noAlign */
    int symmetric;          /* 1 unless -noSym (reg_aladin's default is the symmetric version) */
} nrh_aladin_config;
int nrh_aladin_register(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                        const nrh_aladin_config *cfg, float matrix[16]);
/* reg_aladin<float>::Run (non-symmetric) for n pairs at once: no counterpart in the reference,
 * which registers one pair per process.  Every pair runs the levels, stages and iterations of
 * nrh_aladin_register with the same configuration (cfg->symmetric must be 0), its warps and
 * block matchings launched in turn on one stream and all pairs' LTS fits in one launch (a
 * workgroup per pair); each stage's iteration is captured once as a HIP graph and replayed, so
 * every pair's matrix equals nrh_aladin_register's.  matrices[16 n] receive the reference ->
 * floating affines (row-major); status[n] NRH_OK or the pair's own error (NRH_ERR_NUMERIC: fewer
 * than 2 active blocks, too few correspondences, a singular fit), the other pairs going on.
 * stats may be NULL. */
int nrh_aladin_register_batch(int n, const nrh_geom *refs, const float *const *ref_data, const nrh_geom *flos,
                              const float *const *flo_data, const nrh_aladin_config *cfg, float *matrices,
                              int *status, nrh_batch_stats *stats);

/* ------------------------------------------------------------------ registration server */
/* No counterpart in the reference (one registration per process, reg-apps/reg_aladin.cpp and
 * reg-apps/reg_f3d.cpp).  A request is reg_aladin's affine (nrh_aladin_register with `aladin`)
 * followed by reg_f3d initialised from it (nrh_f3d_register with `f3d`, its affine field set to
 * the Aladin result), NMI or LNCC as f3d.measure says.  Requests are admitted in arrival order
 * when fewer than max_active run and their estimated device memory fits beside the running
 * requests' within hbm_budget_bytes; each runs on its own host thread and HIP streams, so running
 * requests overlap on the device, keep their images, grids and optimiser state resident, and a
 * request's result is the same alone and in any mix.  A failing request (e.g. a flat image:
 * NRH_ERR_NUMERIC) fails alone. */
typedef struct nrh_server nrh_server;
typedef struct nrh_request {
    nrh_geom ref, flo;
    const float *ref_data, *flo_data;  /* copied by nrh_server_submit */
    nrh_aladin_config aladin;
    nrh_f3d_config f3d;                /* f3d.affine is ignored (the Aladin result is used) */
} nrh_request;
typedef struct nrh_result {
    int status;              /* NRH_OK or the request's own error code */
    char error[256];         /* its message */
    float affine[16];        /* the Aladin stage's reference -> floating affine */
    nrh_geom grid;           /* the final control-point grid */
    double similarity;       /* final NMI or LNCC of the F3D stage */
    int iterations;          /* F3D iterations over all levels */
    double submitted, started, finished;  /* seconds since the server was created */
    size_t device_bytes;     /* the admission estimate of the request's device memory */
} nrh_result;
typedef struct nrh_server_stats {
    int pending, active, completed, failed, peak_active;
    size_t peak_bytes, budget_bytes;
} nrh_server_stats;
int nrh_server_create(int device, size_t hbm_budget_bytes, int max_active, nrh_server **out);
/* Returns at once with the request's id (NRH_ERR_NOMEM if the request alone exceeds the budget). */
int nrh_server_submit(nrh_server *s, const nrh_request *req, uint64_t *id);
/* Blocks until request `id` has finished; its result can be taken once.  cpp_out (capacity
 * floats, may be NULL) receives the final control points. */
int nrh_server_wait(nrh_server *s, uint64_t id, nrh_result *res, float *cpp_out, size_t capacity);
int nrh_server_get_stats(nrh_server *s, nrh_server_stats *out);
/* Finishes every submitted request, then frees the server. */
int nrh_server_destroy(nrh_server *s);

/* ------------------------------------------------------------------ file formats
 * The edges of a registration: single-file NIfTI-1 images (.nii / .nii.gz) as
 * reg_io_ReadImageFile / reg_io_WriteImageFile read and write them (reg-io/_reg_ReadWriteImage.cpp:
 * 51-81, 145-170, over niftilib), the control-point grid header of reg_createControlPointGrid
 * (reg-lib/cpu/_reg_localTrans.cpp:21-128) and the affine text matrix of
 * reg_tool_{Read,Write}AffineFile (reg-io/_reg_ReadWriteMatrix.cpp:83-116).  Host code only. */
typedef struct nrh_nifti_info {
    int dim[8];             /* dim[0] = number of axes; after reg_checkAndCorrectDimension */
    float pixdim[8];        /* spacings in mm (micron / metre units converted) */
    int datatype;           /* of the file; data is always returned / written as float32 */
    float scl_slope, scl_inter;  /* slope 1 when the file holds 0 */
    int qform_code, sform_code;
    float quatern[3];       /* quatern_b, c, d */
    float qoffset[3];
    float qfac;             /* -1 or 1 (pixdim[0] of the file) */
    float srow[12];         /* srow_x, srow_y, srow_z */
    float qto_xyz[16];      /* voxel -> mm of the qform as niftilib and reg_checkAndCorrectDimension */
    float sto_xyz[16];      /* set them (row-major; the sform is zero when sform_code <= 0) */
    int intent_code;
    float intent_p[3];
    char intent_name[16];
    float cal_min, cal_max;
    int xyzt_units;
    char descrip[80];
} nrh_nifti_info;
int nrh_nifti_read_info(const char *path, nrh_nifti_info *info);
/* The voxel values converted to float32 (scl_slope / scl_inter not applied); capacity in floats. */
int nrh_nifti_read_data(const char *path, float *data, size_t capacity);
/* The geometry the registrations use: sform if sform_code > 0, else qform (reg-lib/Content.h:31-36). */
int nrh_nifti_geometry(const nrh_nifti_info *info, nrh_geom *geom);
/* The header of the control-point grid reg_f3d writes for a reference image and a grid spacing
 * in mm: dims ceil(n d / s + 3), origin moved by one node in the qform and the sform, intent
 * vector with name "NREG_TRANS" and intent_p1 = CUB_SPLINE_GRID (2). */
int nrh_control_point_grid_info(const nrh_nifti_info *ref, const float spacing[3], nrh_nifti_info *grid);
/* Writes a float32 NIfTI-1 single file (.nii, or gzip-compressed .nii.gz). */
int nrh_nifti_write(const char *path, const nrh_nifti_info *info, const float *data);
/* The 4x4 affine text file: four lines of four numbers ("%.7g"), row-major. */
int nrh_affine_read(const char *path, float m[16]);
int nrh_affine_write(const char *path, const float m[16]);

#ifdef __cplusplus
}
#endif
#endif
