// C-ABI shim over the reference CPU path (test infrastructure only).
//
// Compiled together with the reference's own sources into oracle/_ref/libnrref.so by
// oracle/Makefile.  It lets the tests and the golden-vector generator call the reference
// functions on plain arrays: every image is described by an nr_geom (dims + 4x4 sform,
// row-major) and passed as a planar float array.  Nothing here re-implements reference
// arithmetic; each entry point builds the nifti images and calls the named reference
// function.

#include "_reg_localTrans.h"
#include "_reg_localTrans_regul.h"
#include "_reg_resampling.h"
#include "_reg_nmi.h"
#include "_reg_lncc.h"
#include "_reg_tools.h"
#include "_reg_globalTrans.h"
#include "_reg_blockMatching.h"
#include "_reg_f3d.h"
#include "_reg_aladin.h"
#include "_reg_aladin_sym.h"
#include "F3dContent.h"
#include "Compute.h"
#include "_reg_f3d2.h"
#include "_reg_ReadWriteMatrix.h"
#include <cstring>
#include <vector>
#include <memory>

#include "ref_nifti.hpp"

// The LNCC free functions of reg-lib/cpu/_reg_lncc.cpp (defined there at namespace scope, not
// declared in its header).
double GetSimilarityMeasureValue(const nifti_image *referenceImage, nifti_image *meanImage, nifti_image *sdevImage,
                                 const nifti_image *warpedImage, nifti_image *warpedMeanImage,
                                 nifti_image *warpedSdevImage, const int *referenceMask, int *forwardMask,
                                 const float *kernelStandardDeviation, nifti_image *correlationImage,
                                 const ConvKernelType kernelType, const int referenceTimePoints,
                                 const double *timePointWeights);
void GetVoxelBasedSimilarityMeasureGradient(const nifti_image *referenceImage, nifti_image *meanImage,
                                            nifti_image *sdevImage, const nifti_image *warpedImage,
                                            nifti_image *warpedMeanImage, nifti_image *warpedSdevImage,
                                            const int *referenceMask, int *forwardMask,
                                            const float *kernelStandardDeviation, nifti_image *correlationImage,
                                            const nifti_image *warpedGradient, nifti_image *measureGradient,
                                            const ConvKernelType kernelType, const int currentTimePoint,
                                            const double timePointWeight);

namespace {

using namespace ref_nifti;

thread_local std::string last_error;

// reg_nmi with its per-time-point histogram buffers exposed for comparison.
class RefNmi: public reg_nmi {
public:
    double* Pro() { return this->jointHistogramPro[0]; }
    double* Log() { return this->jointHistogramLog[0]; }
    double* Ent() { return this->entropyValues[0]; }
    int Total() { return this->totalBinNumber[0]; }
};

template<class F> int guarded(F&& f) {
    try { f(); return 0; }
    catch (const std::exception& e) { last_error = e.what(); return 1; }
}

} // namespace

extern "C" {

const char* ref_last_error() { return last_error.c_str(); }

// ---------------------------------------------------------------- control point grid
// reg_createControlPointGrid<float> (cpu/_reg_localTrans.cpp:21-128)
int ref_create_cpp(const nr_geom *ref, const float spacing_mm[3], nr_geom *out, float *data) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, nullptr);
        NiftiImage cpp;
        reg_createControlPointGrid<float>(cpp, refImg, spacing_mm);
        geom_of(cpp, out);
        if (data) std::memcpy(data, cpp->data, cpp.nVoxels() * sizeof(float));
    });
}

// reg_spline_refineControlPointGrid (cpu/_reg_localTrans.cpp:1809-2120); out grid has the
// refined size which the caller obtains by a first call with data_out == nullptr.
int ref_refine_cpp(const nr_geom *ref, const nr_geom *cpp, const float *cppData, nr_geom *out, float *data_out) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, nullptr);
        NiftiImage grid = make_grid(cpp, cppData);
        reg_spline_refineControlPointGrid(grid, refImg);
        geom_of(grid, out);
        if (data_out) std::memcpy(data_out, grid->data, grid.nVoxels() * sizeof(float));
    });
}

// ---------------------------------------------------------------- deformation field
// reg_spline_getDeformationField (cpu/_reg_localTrans.cpp:1457-1538)
int ref_spline_field(const nr_geom *ref, const nr_geom *cpp, const float *cppData, const int *mask,
                     int forceNoLut, float *field) {
    return guarded([&] {
        NiftiImage grid = make_grid(cpp, cppData);
        NiftiImage def = make_vector_image(ref, field);  // masked voxels keep their input value
        std::vector<int> m(mask ? mask : nullptr, mask ? mask + nvox(ref) : nullptr);
        if (!mask) m.assign(nvox(ref), 0);
        reg_spline_getDeformationField(grid, def, m.data(), false, true, forceNoLut != 0);
        std::memcpy(field, def->data, 3 * nvox(ref) * sizeof(float));
    });
}

// reg_affine_getDeformationField (cpu/_reg_globalTrans.cpp:135-175)
int ref_affine_field(const nr_geom *ref, const float affine[16], const int *mask, float *field) {
    return guarded([&] {
        NiftiImage def = make_vector_image(ref, field);
        mat44 a = to_mat44(affine);
        std::vector<int> m(nvox(ref), 0);
        if (mask) std::memcpy(m.data(), mask, nvox(ref) * sizeof(int));
        reg_affine_getDeformationField(&a, def, false, m.data());
        std::memcpy(field, def->data, 3 * nvox(ref) * sizeof(float));
    });
}

// ---------------------------------------------------------------- resampling
// reg_resampleImage (cpu/_reg_resampling.cpp:598-691); warped holds the values kept for masked voxels
int ref_resample(const nr_geom *flo, const float *floData, const nr_geom *ref, const float *field,
                 const int *mask, int interp, float pad, float *warped) {
    return guarded([&] {
        NiftiImage floImg = make_image(flo, floData);
        NiftiImage def = make_vector_image(ref, field);
        NiftiImage war = make_image(ref, warped);
        std::vector<int> m(nvox(ref), 0);
        if (mask) std::memcpy(m.data(), mask, nvox(ref) * sizeof(int));
        reg_resampleImage(floImg, war, def, m.data(), interp, pad);
        std::memcpy(warped, war->data, nvox(ref) * sizeof(float));
    });
}

// reg_getImageGradient (cpu/_reg_resampling.cpp:2269-2384)
int ref_image_gradient(const nr_geom *flo, const float *floData, const nr_geom *ref, const float *field,
                       const int *mask, int interp, float pad, float *grad) {
    return guarded([&] {
        NiftiImage floImg = make_image(flo, floData);
        NiftiImage def = make_vector_image(ref, field);
        NiftiImage g = make_vector_image(ref, nullptr);
        std::vector<int> m(nvox(ref), 0);
        if (mask) std::memcpy(m.data(), mask, nvox(ref) * sizeof(int));
        reg_getImageGradient(floImg, g, def, m.data(), interp, pad, 0);
        std::memcpy(grad, g->data, 3 * nvox(ref) * sizeof(float));
    });
}

// ---------------------------------------------------------------- NMI
// reg_intensityRescale (cpu/_reg_tools.cpp:93-208)
int ref_intensity_rescale(const nr_geom *g, float *data, float newMin, float newMax) {
    return guarded([&] {
        NiftiImage img = make_image(g, data);
        reg_intensityRescale(img, 0, newMin, newMax);
        std::memcpy(data, img->data, nvox(g) * sizeof(float));
    });
}

// reg_nmi value (cpu/_reg_nmi.cpp:160-398) on already-rescaled images.
// histPro / histLog: totalBins doubles each (bins_r*bins_f + bins_r + bins_f); ent: 4 doubles.
int ref_nmi_value(const nr_geom *ref, const float *refData, const float *warped, const int *mask,
                  int refBins, int floBins, int approx, double *value, double *histPro, double *histLog,
                  double *ent) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, refData);
        NiftiImage floImg = make_image(ref, refData);
        NiftiImage war = make_image(ref, warped);
        NiftiImage wg = make_vector_image(ref, nullptr);
        NiftiImage vg = make_vector_image(ref, nullptr);
        std::vector<int> m(nvox(ref), 0);
        if (mask) std::memcpy(m.data(), mask, nvox(ref) * sizeof(int));
        // Bypass the in-place rescale of InitialiseMeasure: the caller passes rescaled data,
        // so the measure is initialised on images whose range is already [2, bins-3].
        RefNmi nmi;
        nmi.SetTimePointWeight(0, 1.0);
        nmi.SetReferenceBinNumber(refBins, 0);
        nmi.SetFloatingBinNumber(floBins, 0);
        if (!approx) nmi.DoNotApproximatePw();
        nmi.InitialiseMeasure(refImg, floImg, m.data(), war, wg, vg);
        // InitialiseMeasure rescaled refImg; restore the caller's values
        std::memcpy(refImg->data, refData, nvox(ref) * sizeof(float));
        *value = nmi.GetSimilarityMeasureValue();
        if (histPro) std::memcpy(histPro, nmi.Pro(), nmi.Total() * sizeof(double));
        if (histLog) std::memcpy(histLog, nmi.Log(), nmi.Total() * sizeof(double));
        if (ent) std::memcpy(ent, nmi.Ent(), 4 * sizeof(double));
    });
}

// NMI value + voxel-based gradient (cpu/_reg_nmi.cpp:501-638).  voxelGrad is accumulated (+=).
int ref_nmi_gradient(const nr_geom *ref, const float *refData, const float *warped, const float *warpedGrad,
                     const int *mask, int refBins, int floBins, double *value, float *voxelGrad) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, refData);
        NiftiImage floImg = make_image(ref, refData);
        NiftiImage war = make_image(ref, warped);
        NiftiImage wg = make_vector_image(ref, warpedGrad);
        NiftiImage vg = make_vector_image(ref, voxelGrad);
        std::vector<int> m(nvox(ref), 0);
        if (mask) std::memcpy(m.data(), mask, nvox(ref) * sizeof(int));
        reg_nmi nmi;
        nmi.SetTimePointWeight(0, 1.0);
        nmi.SetReferenceBinNumber(refBins, 0);
        nmi.SetFloatingBinNumber(floBins, 0);
        nmi.InitialiseMeasure(refImg, floImg, m.data(), war, wg, vg);
        std::memcpy(refImg->data, refData, nvox(ref) * sizeof(float));
        *value = nmi.GetSimilarityMeasureValue();
        nmi.GetVoxelBasedSimilarityMeasureGradient(0);
        std::memcpy(voxelGrad, vg->data, 3 * nvox(ref) * sizeof(float));
    });
}

// ---------------------------------------------------------------- convolution / node gradient
// reg_tools_kernelConvolution (cpu/_reg_tools.cpp:832-1341) on an image with nt volumes.
// kernel: 0 mean, 1 linear, 2 gaussian, 3 cubic (ConvKernelType order).
int ref_kernel_convolution(const nr_geom *g, int nt, float *data, const float *sigma, int kernel,
                           const int *mask, const int axes[3]) {
    return guarded([&] {
        NiftiImage img = make_image(g, data, nt, 1);
        std::unique_ptr<bool[]> tp(new bool[nt]);
        for (int i = 0; i < nt; ++i) tp[i] = true;
        bool ax[3] = { axes[0] != 0, axes[1] != 0, axes[2] != 0 };
        reg_tools_kernelConvolution(img, sigma, static_cast<ConvKernelType>(kernel), mask, tp.get(), ax);
        std::memcpy(data, img->data, nvox(g) * nt * sizeof(float));
    });
}

// Compute::ConvolveImage (Compute.cpp:230-264) followed by Compute::VoxelCentricToNodeCentric
// (Compute.cpp:266-274 -> cpu/_reg_localTrans.cpp:1540-1705).  voxelGrad is convolved in place.
int ref_node_gradient(const nr_geom *ref, float *voxelGrad, const nr_geom *cpp, const nr_geom *flo,
                      float weight, float *nodeGrad) {
    return guarded([&] {
        NiftiImage vg = make_vector_image(ref, voxelGrad);
        NiftiImage grid = make_grid(cpp, nullptr);
        NiftiImage floImg = make_image(flo, nullptr);
        const int axesOrder[3][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
        const float sp[3] = { grid->dx, grid->dy, grid->dz };
        for (int a = 0; a < 3; ++a) {
            float s[3] = { sp[a], sp[a], sp[a] };
            bool ax[3] = { axesOrder[a][0] != 0, axesOrder[a][1] != 0, axesOrder[a][2] != 0 };
            reg_tools_kernelConvolution(vg, s, ConvKernelType::Cubic, nullptr, nullptr, ax);
        }
        std::memcpy(voxelGrad, vg->data, 3 * nvox(ref) * sizeof(float));
        NiftiImage ng = make_grid(cpp, nullptr);
        mat44 *reorient = Content::GetIJKMatrix(*floImg);
        reg_voxelCentricToNodeCentric(ng, vg, weight, false, reorient);
        std::memcpy(nodeGrad, ng->data, 3 * nvox(cpp) * sizeof(float));
    });
}

// ---------------------------------------------------------------- bending energy
// reg_spline_approxBendingEnergy (cpu/_reg_localTrans_regul.cpp:73-177)
int ref_be_value(const nr_geom *cpp, const float *cppData, double *value) {
    return guarded([&] {
        NiftiImage grid = make_grid(cpp, cppData);
        *value = reg_spline_approxBendingEnergy(grid);
    });
}

// reg_spline_approxBendingEnergyGradient (cpu/_reg_localTrans_regul.cpp:287-482); grad is += .
// The reference converts the grid to displacement and back in place; the round-tripped grid is
// returned in cppOut when non-null.
int ref_be_gradient(const nr_geom *cpp, const float *cppData, float weight, float *grad, float *cppOut) {
    return guarded([&] {
        NiftiImage grid = make_grid(cpp, cppData);
        NiftiImage g = make_grid(cpp, grad);
        reg_spline_approxBendingEnergyGradient(grid, g, weight);
        std::memcpy(grad, g->data, 3 * nvox(cpp) * sizeof(float));
        if (cppOut) std::memcpy(cppOut, grid->data, 3 * nvox(cpp) * sizeof(float));
    });
}

// reg_spline_approxLinearEnergy / Gradient (cpu/_reg_localTrans_regul.cpp:560-857)
int ref_le_value(const nr_geom *cpp, const float *cppData, double *value) {
    return guarded([&] {
        NiftiImage grid = make_grid(cpp, cppData);
        *value = reg_spline_approxLinearEnergy(grid);
    });
}
int ref_le_gradient(const nr_geom *cpp, const float *cppData, float weight, float *grad) {
    return guarded([&] {
        NiftiImage grid = make_grid(cpp, cppData);
        NiftiImage g = make_grid(cpp, grad);
        reg_spline_approxLinearEnergyGradient(grid, g, weight);
        std::memcpy(grad, g->data, 3 * nvox(cpp) * sizeof(float));
    });
}

// ---------------------------------------------------------------- pyramid
// reg_downsampleImage<float> (cpu/_reg_tools.cpp:1343-1467), type 1 = smoothed + decimated.
// First call with out data == nullptr to get the output geometry.
int ref_downsample(const nr_geom *g, const float *data, nr_geom *out, float *outData) {
    return guarded([&] {
        NiftiImage img = make_image(g, data);
        bool axes[8] = { false, true, true, true, false, false, false, false };
        reg_downsampleImage<float>(img, 1, axes);
        geom_of(img, out);
        if (outData) std::memcpy(outData, img->data, img.nVoxels() * sizeof(float));
    });
}


// ---------------------------------------------------------------- velocity fields (F3D2)
static NiftiImage make_velocity_grid(const nr_geom *cpp, const float *data, int steps) {
    NiftiImage grid = make_grid(cpp, data);
    grid->intent_p1 = SPLINE_VEL_GRID;
    grid->intent_p2 = (float)steps;
    return grid;
}

// reg_spline_getDefFieldFromVelocityGrid (cpu/_reg_localTrans.cpp:3583-3613); *steps: intent_p2
int ref_vel_def_field(const nr_geom *ref, const nr_geom *cpp, const float *vel, int *steps, int update,
                      float *field) {
    return guarded([&] {
        NiftiImage grid = make_velocity_grid(cpp, vel, *steps);
        NiftiImage def = make_vector_image(ref, nullptr);
        reg_spline_getDefFieldFromVelocityGrid(grid, def, update != 0);
        *steps = (int)grid->intent_p2;
        std::memcpy(field, def->data, 3 * nvox(ref) * sizeof(float));
    });
}

// reg_spline_getIntermediateDefFieldFromVelGrid (cpu/_reg_localTrans.cpp:3616-3683)
int ref_vel_intermediate(const nr_geom *ref, const nr_geom *cpp, const float *vel, int steps, float *fields) {
    return guarded([&] {
        NiftiImage grid = make_velocity_grid(cpp, vel, steps);
        const int sq = steps < 0 ? -steps : steps;
        std::vector<NiftiImage> defs(sq + 1);
        for (auto& d : defs) d = make_vector_image(ref, nullptr);
        reg_spline_getIntermediateDefFieldFromVelGrid(grid, defs.data());
        for (int k = 0; k <= sq; ++k)
            std::memcpy(fields + (size_t)k * 3 * nvox(ref), defs[k]->data, 3 * nvox(ref) * sizeof(float));
    });
}

// reg_defField_compose (cpu/_reg_localTrans.cpp:2460-2476), no mask; upd in place
int ref_def_field_compose(const nr_geom *g, const float *def, float *upd) {
    return guarded([&] {
        NiftiImage d = make_vector_image(g, def);
        NiftiImage u = make_vector_image(g, upd);
        reg_defField_compose(d, u, nullptr);
        std::memcpy(upd, u->data, 3 * nvox(g) * sizeof(float));
    });
}

// reg_resampleGradient (cpu/_reg_resampling.cpp:1769-1797), linear
int ref_resample_gradient(const nr_geom *g, const float *grad, const float *def, float pad, float *out) {
    return guarded([&] {
        NiftiImage gi = make_vector_image(g, grad);
        NiftiImage di = make_vector_image(g, def);
        NiftiImage oi = make_vector_image(g, nullptr);
        reg_resampleGradient(gi, oi, di, 1, pad);
        std::memcpy(out, oi->data, 3 * nvox(g) * sizeof(float));
    });
}

// Compute::ExponentiateGradient (Compute.cpp:283-326) on two F3dContent objects: the forward
// one (reference -> floating) holding `grad` as its voxel-based gradient, the backward one
// (floating -> reference) holding the backward velocity grid.
int ref_exponentiate_gradient(const nr_geom *ref, const nr_geom *flo, const nr_geom *cpp, const float *velFw,
                              const float *velBw, int steps, float *grad) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, nullptr), floImg = make_image(flo, nullptr);
        NiftiImage gridFw = make_velocity_grid(cpp, velFw, steps), gridBw = make_velocity_grid(cpp, velBw, steps);
        F3dContent con(refImg, floImg, gridFw), conBw(floImg, refImg, gridBw);
        NiftiImage& vg = con.GetVoxelBasedMeasureGradient();
        std::memcpy(vg->data, grad, 3 * nvox(ref) * sizeof(float));
        Compute(con).ExponentiateGradient(conBw);
        std::memcpy(grad, vg->data, 3 * nvox(ref) * sizeof(float));
    });
}

// Compute::UpdateVelocityField (Compute.cpp:336-348) and Compute::SymmetriseVelocityFields
// (Compute.cpp:357-391) on a forward / backward content pair.  mode 0: update the forward grid
// with `grad` (scale, axes); mode 1: symmetrise both grids.
int ref_velocity_update(const nr_geom *ref, const nr_geom *flo, const nr_geom *cpp, float *velFw, float *velBw,
                        const float *grad, float scale, const int axes[3], int mode) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, nullptr), floImg = make_image(flo, nullptr);
        NiftiImage gridFw = make_velocity_grid(cpp, velFw, 6), gridBw = make_velocity_grid(cpp, velBw, 6);
        F3dContent con(refImg, floImg, gridFw), conBw(floImg, refImg, gridBw);
        Compute compute(con);
        if (mode == 0) {
            NiftiImage& tg = con.GetTransformationGradient();
            std::memcpy(tg->data, grad, 3 * nvox(cpp) * sizeof(float));
            compute.UpdateVelocityField(scale, axes[0] != 0, axes[1] != 0, axes[2] != 0);
        } else {
            compute.SymmetriseVelocityFields(conBw);
        }
        std::memcpy(velFw, con.GetControlPointGrid()->data, 3 * nvox(cpp) * sizeof(float));
        std::memcpy(velBw, conBw.GetControlPointGrid()->data, 3 * nvox(cpp) * sizeof(float));
    });
}

// reg_createSymmetricControlPointGrids<float> without an affine (cpu/_reg_localTrans.cpp:132-358)
int ref_symmetric_grids(const nr_geom *ref, const nr_geom *flo, const float spacing[3], nr_geom *out, float *fw,
                        float *bw, int capacity) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, nullptr), floImg = make_image(flo, nullptr);
        NiftiImage gFw, gBw;
        reg_createSymmetricControlPointGrids<float>(gFw, gBw, refImg, floImg, nullptr, spacing);
        geom_of(gFw, out);
        if ((int)gFw.nVoxels() <= capacity) {
            if (fw) std::memcpy(fw, gFw->data, gFw.nVoxels() * sizeof(float));
            if (bw) std::memcpy(bw, gBw->data, gBw.nVoxels() * sizeof(float));
        }
    });
}


// ---------------------------------------------------------------- velocity grids with affine parts
// A grid's NIfTI extensions 0 and 1 hold mat44 affines (reg-lib/cpu/_reg_localTrans.cpp:320-353).
static void attach_affines(NiftiImage& grid, const float *ext0, const float *ext1) {
    for (const float *e : { ext0, ext1 }) {
        if (!e) break;
        mat44 m = to_mat44(e);
        nifti_add_extension(grid, reinterpret_cast<const char *>(&m), (int)sizeof(mat44), NIFTI_ECODE_IGNORE);
    }
}

// reg_createSymmetricControlPointGrids<float> with an affine: geometry, positions and the forward
// / backward grids' extension 0 (*count: the number of extensions, 0 or 2).
int ref_symmetric_grids_aff(const nr_geom *ref, const nr_geom *flo, const float spacing[3], const float *affine,
                            nr_geom *out, float *fw, float *bw, int capacity, float *extFw, float *extBw, int *count) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, nullptr), floImg = make_image(flo, nullptr);
        NiftiImage gFw, gBw;
        mat44 a = to_mat44(affine);
        reg_createSymmetricControlPointGrids<float>(gFw, gBw, refImg, floImg, &a, spacing);
        geom_of(gFw, out);
        if ((int)gFw.nVoxels() <= capacity) {
            if (fw) std::memcpy(fw, gFw->data, gFw.nVoxels() * sizeof(float));
            if (bw) std::memcpy(bw, gBw->data, gBw.nVoxels() * sizeof(float));
        }
        *count = gFw->num_ext;
        if (gFw->num_ext > 0) {
            from_mat44(*reinterpret_cast<mat44 *>(gFw->ext_list[0].edata), extFw);
            from_mat44(*reinterpret_cast<mat44 *>(gBw->ext_list[0].edata), extBw);
        }
    });
}

int ref_vel_def_field_aff(const nr_geom *ref, const nr_geom *cpp, const float *vel, const float *ext0,
                          const float *ext1, int *steps, int update, float *field) {
    return guarded([&] {
        NiftiImage grid = make_velocity_grid(cpp, vel, *steps);
        attach_affines(grid, ext0, ext1);
        NiftiImage def = make_vector_image(ref, nullptr);
        reg_spline_getDefFieldFromVelocityGrid(grid, def, update != 0);
        *steps = (int)grid->intent_p2;
        std::memcpy(field, def->data, 3 * nvox(ref) * sizeof(float));
    });
}

int ref_vel_intermediate_aff(const nr_geom *ref, const nr_geom *cpp, const float *vel, const float *ext0,
                             const float *ext1, int steps, float *fields) {
    return guarded([&] {
        NiftiImage grid = make_velocity_grid(cpp, vel, steps);
        attach_affines(grid, ext0, ext1);
        const int sq = steps < 0 ? -steps : steps;
        std::vector<NiftiImage> defs(sq + 1);
        for (auto& d : defs) d = make_vector_image(ref, nullptr);
        reg_spline_getIntermediateDefFieldFromVelGrid(grid, defs.data());
        for (int k = 0; k <= sq; ++k)
            std::memcpy(fields + (size_t)k * 3 * nvox(ref), defs[k]->data, 3 * nvox(ref) * sizeof(float));
    });
}

// Compute::ExponentiateGradient with the backward grid's extensions and the backward content's
// transformation matrix.
int ref_exponentiate_gradient_aff(const nr_geom *ref, const nr_geom *flo, const nr_geom *cpp, const float *velFw,
                                  const float *velBw, const float *ext0, const float *ext1, const float *transBw,
                                  int steps, float *grad) {
    return guarded([&] {
        NiftiImage refImg = make_image(ref, nullptr), floImg = make_image(flo, nullptr);
        NiftiImage gridFw = make_velocity_grid(cpp, velFw, steps), gridBw = make_velocity_grid(cpp, velBw, steps);
        attach_affines(gridBw, ext0, ext1);
        mat44 t = transBw ? to_mat44(transBw) : mat44{};
        F3dContent con(refImg, floImg, gridFw), conBw(floImg, refImg, gridBw, nullptr, nullptr, transBw ? &t : nullptr);
        NiftiImage& vg = con.GetVoxelBasedMeasureGradient();
        std::memcpy(vg->data, grad, 3 * nvox(ref) * sizeof(float));
        Compute(con).ExponentiateGradient(conBw);
        std::memcpy(grad, vg->data, 3 * nvox(ref) * sizeof(float));
    });
}

// reg_voxelCentricToNodeCentric (no convolution) on a grid with extension 0; reorientation from
// the floating image's mm -> voxel matrix as Compute::VoxelCentricToNodeCentric passes it.
int ref_voxel_to_node_ext(const nr_geom *cpp, const nr_geom *ref, const nr_geom *flo, const float *voxelGrad,
                          const float *ext0, float weight, float *nodeGrad) {
    return guarded([&] {
        NiftiImage vg = make_vector_image(ref, voxelGrad);
        NiftiImage floImg = make_image(flo, nullptr);
        NiftiImage ng = make_grid(cpp, nullptr);
        attach_affines(ng, ext0, nullptr);
        reg_voxelCentricToNodeCentric(ng, vg, weight, false, Content::GetIJKMatrix(*floImg));
        std::memcpy(nodeGrad, ng->data, 3 * nvox(cpp) * sizeof(float));
    });
}

} // extern "C"

// ---------------------------------------------------------------- full F3D through reg_f3d
namespace {
class RefF3d: public reg_f3d<float> {
public:
    RefF3d(): reg_f3d<float>(1, 1) {}
    // Prepare level `level` the way reg_base::Run does before its optimisation loop.
    void Prepare(int level) {
        this->Initialise();
        this->maxIterationNumber *= 1;
        // the grid of `level` is the coarsest grid refined once per level before it, as in Run()
        for (int l = 0; l < level; ++l) {
            this->InitCurrentLevel(l);
            this->DeinitCurrentLevel(l);
        }
        this->InitCurrentLevel(level);
        this->InitialiseSimilarity();
        this->SetOptimiser();
    }
    double Value() { return this->GetObjectiveFunctionValue(); }
    void Gradient() { this->GetObjectiveFunctionGradient(); }
    double NormGrad() { return this->NormaliseGradient(); }
    F3dContent& Con() { return dynamic_cast<F3dContent&>(*this->con); }
    NiftiImage& Grid() { return this->controlPointGrid; }
    double WMeasure() { return this->currentWMeasure; }
    double WBE() { return this->currentWBE; }
    double WLE() { return this->currentWLE; }
    float SimWeight() { return this->similarityWeight; }
    float BEWeight() { return this->bendingEnergyWeight; }
    float LEWeight() { return this->linearEnergyWeight; }
    size_t Iterations() { return this->optimiser ? this->optimiser->GetCurrentIterationNumber() : 0; }
};

struct F3dHandle {
    std::unique_ptr<RefF3d> reg;
    nr_geom refG, floG;
};
} // namespace

// Runs print the reference's per-iteration objective lines when set (diagnostics).
static bool g_verbose = false;

extern "C" {

int ref_set_verbose(int v) {
    g_verbose = v != 0;
    return 0;
}

typedef struct {
    int bins;            // user bin count (reference adds 4)
    float be;            // bending energy weight
    float le;            // linear energy weight
    float spacing;       // control point spacing; negative = voxels
    int levels;          // pyramid levels
    int levels_to_perform;
    int maxit;           // final-level iteration budget
    int use_pyramid;
    float padding;       // warped padding value
    int measure;         // 0 NMI, 1 LNCC
    float lncc_sigma;    // LNCC kernel standard deviation (< 0: voxels)
    const float *affine; // input affine (reference -> floating), NULL for none
} ref_f3d_params;

// The platform is chosen first, as reg_f3d's command line does (reg-apps/reg_f3d.cpp:287): the
// measures are created by the platform's measure creator when they are configured.
static void configure(reg_f3d<float>& r, const ref_f3d_params *p, PlatformType platform = PlatformType::Cpu) {
    r.SetPlatformType(platform);
    if (p->measure == 1) {
        r.UseLNCC(0, p->lncc_sigma);
    } else {
        r.UseNMISetReferenceBinNumber(0, p->bins);
        r.UseNMISetFloatingBinNumber(0, p->bins);
    }
    r.SetBendingEnergyWeight(p->be);
    r.SetLinearEnergyWeight(p->le);
    r.SetJacobianLogWeight(0);
    r.SetSpacing(0, p->spacing);
    r.SetLevelNumber(p->levels);
    r.SetLevelToPerform(p->levels_to_perform);
    r.SetMaximalIterationNumber(p->maxit);
    if (!p->use_pyramid) r.DoNotUsePyramidalApproach();
    r.SetWarpedPaddingValue(p->padding);
    if (p->affine) r.SetAffineTransformation(to_mat44(p->affine));
    if (!g_verbose) r.DoNotPrintOutInformation();
}

// Create a reg_f3d<float>, run Initialise / InitCurrentLevel(level) / InitialiseSimilarity /
// SetOptimiser so that single objective+gradient evaluations can be timed and compared.
void* ref_f3d_create(const nr_geom *ref, const float *refData, const nr_geom *flo, const float *floData,
                     const ref_f3d_params *p, int level) {
    F3dHandle *h = new F3dHandle;
    int rc = guarded([&] {
        h->reg.reset(new RefF3d);
        h->refG = *ref; h->floG = *flo;
        configure(*h->reg, p);
        h->reg->SetReferenceImage(make_image(ref, refData));
        h->reg->SetFloatingImage(make_image(flo, floData));
        h->reg->Prepare(level);
    });
    if (rc) { delete h; return nullptr; }
    return h;
}

void ref_f3d_destroy(void *hp) { delete static_cast<F3dHandle*>(hp); }

// Geometry of the current level's grid and images.
int ref_f3d_level_geom(void *hp, nr_geom *cpp, nr_geom *ref, nr_geom *flo) {
    F3dHandle *h = static_cast<F3dHandle*>(hp);
    return guarded([&] {
        F3dContent& c = h->reg->Con();
        geom_of(c.F3dContent::GetControlPointGrid(), cpp);
        geom_of(c.Content::GetReference(), ref);
        geom_of(c.Content::GetFloating(), flo);
    });
}

// Copy out the current level's (rescaled) reference / floating images.
int ref_f3d_level_images(void *hp, float *refData, float *floData) {
    F3dHandle *h = static_cast<F3dHandle*>(hp);
    return guarded([&] {
        F3dContent& c = h->reg->Con();
        NiftiImage& r = c.Content::GetReference();
        NiftiImage& f = c.Content::GetFloating();
        if (refData) std::memcpy(refData, r->data, r.nVoxelsPerVolume() * sizeof(float));
        if (floData) std::memcpy(floData, f->data, f.nVoxelsPerVolume() * sizeof(float));
    });
}

int ref_f3d_set_cpp(void *hp, const float *cppData) {
    F3dHandle *h = static_cast<F3dHandle*>(hp);
    return guarded([&] {
        NiftiImage& g = h->reg->Con().F3dContent::GetControlPointGrid();
        std::memcpy(g->data, cppData, g.nVoxels() * sizeof(float));
    });
}

int ref_f3d_get_cpp(void *hp, float *cppData) {
    F3dHandle *h = static_cast<F3dHandle*>(hp);
    return guarded([&] {
        NiftiImage& g = h->reg->Con().F3dContent::GetControlPointGrid();
        std::memcpy(cppData, g->data, g.nVoxels() * sizeof(float));
    });
}

// reg_f3d::GetObjectiveFunctionValue (_reg_f3d.cpp:443-463).  terms: wMeasure, wBE, wLE.
int ref_f3d_value(void *hp, double *value, double *terms) {
    F3dHandle *h = static_cast<F3dHandle*>(hp);
    return guarded([&] {
        *value = h->reg->Value();
        if (terms) { terms[0] = h->reg->WMeasure(); terms[1] = h->reg->WBE(); terms[2] = h->reg->WLE(); }
    });
}

// reg_f3d::GetObjectiveFunctionGradient (_reg_f3d.cpp:553-577), optionally followed by
// reg_f3d::NormaliseGradient (_reg_f3d.cpp:395-411).  grad receives the transformation gradient.
int ref_f3d_gradient(void *hp, int normalise, double *maxLength, float *grad, float *voxelGrad) {
    F3dHandle *h = static_cast<F3dHandle*>(hp);
    return guarded([&] {
        h->reg->Gradient();
        if (normalise) { double m = h->reg->NormGrad(); if (maxLength) *maxLength = m; }
        F3dContent& c = h->reg->Con();
        NiftiImage& tg = c.GetTransformationGradient();
        if (grad) std::memcpy(grad, tg->data, tg.nVoxels() * sizeof(float));
        if (voxelGrad) {
            NiftiImage& vg = c.GetVoxelBasedMeasureGradient();
            std::memcpy(voxelGrad, vg->data, vg.nVoxels() * sizeof(float));
        }
    });
}

int ref_f3d_weights(void *hp, float *w) {
    F3dHandle *h = static_cast<F3dHandle*>(hp);
    w[0] = h->reg->SimWeight(); w[1] = h->reg->BEWeight(); w[2] = h->reg->LEWeight();
    return 0;
}

// Full registration through reg_base::Run (_reg_base.cpp:727-819) on the given platform
// (PlatformType value; 3 is the HIP platform of the patched build).  Output: final-level grid.
// Input images for Run(); named, so that a verbose run can print their names.
static NiftiImage run_image(const nr_geom *g, const float *data, const char *name) {
    NiftiImage img = make_image(g, data);
    img->fname = nifti_strdup(name);
    img->iname = nifti_strdup(name);
    return img;
}

int ref_f3d_run_on(int platform, const nr_geom *ref, const float *refData, const nr_geom *flo, const float *floData,
                   const ref_f3d_params *p, nr_geom *cppOut, float *cppData, int cppCapacity) {
    return guarded([&] {
        RefF3d r;
        configure(r, p, static_cast<PlatformType>(platform));
        r.SetReferenceImage(run_image(ref, refData, "reference"));
        r.SetFloatingImage(run_image(flo, floData, "floating"));
        r.Run();
        NiftiImage g = r.GetControlPointPositionImage();
        geom_of(g, cppOut);
        if (cppData && (int)g.nVoxels() <= cppCapacity)
            std::memcpy(cppData, g->data, g.nVoxels() * sizeof(float));
    });
}

// Full registration through reg_base::Run (_reg_base.cpp:727-819).  Output: final-level grid.
int ref_f3d_run(const nr_geom *ref, const float *refData, const nr_geom *flo, const float *floData,
                const ref_f3d_params *p, nr_geom *cppOut, float *cppData, int cppCapacity) {
    return guarded([&] {
        RefF3d r;
        configure(r, p);
        r.SetReferenceImage(make_image(ref, refData));
        r.SetFloatingImage(make_image(flo, floData));
        r.Run();
        NiftiImage g = r.GetControlPointPositionImage();
        geom_of(g, cppOut);
        if (cppData && (int)g.nVoxels() <= cppCapacity)
            std::memcpy(cppData, g->data, g.nVoxels() * sizeof(float));
    });
}

} // extern "C"

// ---------------------------------------------------------------- symmetric F3D through reg_f3d2
namespace {
class RefF3d2: public reg_f3d2<float> {
public:
    RefF3d2(): reg_f3d2<float>(1, 1) {}
    void Prepare(int level) {
        this->Initialise();
        for (int l = 0; l < level; ++l) {
            this->InitCurrentLevel(l);
            this->DeinitCurrentLevel(l);
        }
        this->InitCurrentLevel(level);
        this->InitialiseSimilarity();
        this->SetOptimiser();
    }
    double Value() { return this->GetObjectiveFunctionValue(); }
    void Gradient() { this->GetObjectiveFunctionGradient(); }
    double NormGrad() { return this->NormaliseGradient(); }
    F3dContent& Con() { return dynamic_cast<F3dContent&>(*this->con); }
    F3dContent& ConBw() { return *this->conBw; }
    double WMeasure() { return this->currentWMeasure; }
    double WBE() { return this->currentWBE; }
    double WLE() { return this->currentWLE; }
    float SimWeight() { return this->similarityWeight; }
    float BEWeight() { return this->bendingEnergyWeight; }
    float LEWeight() { return this->linearEnergyWeight; }
};

struct F3d2Handle {
    std::unique_ptr<RefF3d2> reg;
};
} // namespace

extern "C" {

// reg_f3d2<float> prepared like reg_f3d (Initialise, InitCurrentLevel, InitialiseSimilarity,
// SetOptimiser) so that single symmetric objective / gradient evaluations can be timed and
// compared.
void* ref_f3d2_create(const nr_geom *ref, const float *refData, const nr_geom *flo, const float *floData,
                      const ref_f3d_params *p, int level) {
    F3d2Handle *h = new F3d2Handle;
    int rc = guarded([&] {
        h->reg.reset(new RefF3d2);
        configure(*h->reg, p);
        h->reg->SetReferenceImage(make_image(ref, refData));
        h->reg->SetFloatingImage(make_image(flo, floData));
        h->reg->Prepare(level);
    });
    if (rc) { delete h; return nullptr; }
    return h;
}

void ref_f3d2_destroy(void *hp) { delete static_cast<F3d2Handle*>(hp); }

int ref_f3d2_level_geom(void *hp, nr_geom *cpp, nr_geom *ref, nr_geom *flo) {
    F3d2Handle *h = static_cast<F3d2Handle*>(hp);
    return guarded([&] {
        F3dContent& c = h->reg->Con();
        geom_of(c.F3dContent::GetControlPointGrid(), cpp);
        geom_of(c.Content::GetReference(), ref);
        geom_of(c.Content::GetFloating(), flo);
    });
}

int ref_f3d2_level_images(void *hp, float *refData, float *floData) {
    F3d2Handle *h = static_cast<F3d2Handle*>(hp);
    return guarded([&] {
        F3dContent& c = h->reg->Con();
        NiftiImage& r = c.Content::GetReference();
        NiftiImage& f = c.Content::GetFloating();
        if (refData) std::memcpy(refData, r->data, r.nVoxelsPerVolume() * sizeof(float));
        if (floData) std::memcpy(floData, f->data, f.nVoxelsPerVolume() * sizeof(float));
    });
}

// Forward / backward grids (positions) and their intent_p2 step numbers.
int ref_f3d2_set_cpp(void *hp, const float *fw, const float *bw) {
    F3d2Handle *h = static_cast<F3d2Handle*>(hp);
    return guarded([&] {
        NiftiImage& g = h->reg->Con().F3dContent::GetControlPointGrid();
        NiftiImage& gb = h->reg->ConBw().F3dContent::GetControlPointGrid();
        if (fw) std::memcpy(g->data, fw, g.nVoxels() * sizeof(float));
        if (bw) std::memcpy(gb->data, bw, gb.nVoxels() * sizeof(float));
    });
}

int ref_f3d2_get_cpp(void *hp, float *fw, float *bw, int *steps) {
    F3d2Handle *h = static_cast<F3d2Handle*>(hp);
    return guarded([&] {
        NiftiImage& g = h->reg->Con().F3dContent::GetControlPointGrid();
        NiftiImage& gb = h->reg->ConBw().F3dContent::GetControlPointGrid();
        if (fw) std::memcpy(fw, g->data, g.nVoxels() * sizeof(float));
        if (bw) std::memcpy(bw, gb->data, gb.nVoxels() * sizeof(float));
        if (steps) { steps[0] = (int)g->intent_p2; steps[1] = (int)gb->intent_p2; }
    });
}

// reg_f3d2::GetObjectiveFunctionValue (_reg_f3d2.cpp:540-560).  terms: wMeasure, wBE, wLE.
int ref_f3d2_value(void *hp, double *value, double *terms) {
    F3d2Handle *h = static_cast<F3d2Handle*>(hp);
    return guarded([&] {
        *value = h->reg->Value();
        if (terms) { terms[0] = h->reg->WMeasure(); terms[1] = h->reg->WBE(); terms[2] = h->reg->WLE(); }
    });
}

// reg_f3d2::GetObjectiveFunctionGradient (_reg_f3d2.cpp:417-440), optionally followed by
// reg_f3d2::NormaliseGradient (:373-395).  Transformation gradients of both directions.
int ref_f3d2_gradient(void *hp, int normalise, double *maxLength, float *gradFw, float *gradBw) {
    F3d2Handle *h = static_cast<F3d2Handle*>(hp);
    return guarded([&] {
        h->reg->Gradient();
        if (normalise) { double m = h->reg->NormGrad(); if (maxLength) *maxLength = m; }
        NiftiImage& tg = h->reg->Con().GetTransformationGradient();
        NiftiImage& tb = h->reg->ConBw().GetTransformationGradient();
        if (gradFw) std::memcpy(gradFw, tg->data, tg.nVoxels() * sizeof(float));
        if (gradBw) std::memcpy(gradBw, tb->data, tb.nVoxels() * sizeof(float));
    });
}

// Full symmetric registration through reg_f3d2's run (reg_base::Run with the reg_f3d2 overrides)
// on the given platform.  Output: the final-level grid geometry, both velocity grids and their
// intent_p2.
int ref_f3d2_run_on(int platform, const nr_geom *ref, const float *refData, const nr_geom *flo, const float *floData,
                    const ref_f3d_params *p, nr_geom *cppOut, float *fw, float *bw, int capacity, int *steps) {
    return guarded([&] {
        RefF3d2 r;
        configure(r, p, static_cast<PlatformType>(platform));
        r.SetReferenceImage(run_image(ref, refData, "reference"));
        r.SetFloatingImage(run_image(flo, floData, "floating"));
        r.Run();
        NiftiImage g = r.GetControlPointPositionImage();
        NiftiImage gb = r.GetBackwardControlPointPositionImage();
        geom_of(g, cppOut);
        if ((int)g.nVoxels() <= capacity) {
            if (fw) std::memcpy(fw, g->data, g.nVoxels() * sizeof(float));
            if (bw) std::memcpy(bw, gb->data, gb.nVoxels() * sizeof(float));
        }
        if (steps) { steps[0] = (int)g->intent_p2; steps[1] = (int)gb->intent_p2; }
    });
}

int ref_f3d2_run(const nr_geom *ref, const float *refData, const nr_geom *flo, const float *floData,
                 const ref_f3d_params *p, nr_geom *cppOut, float *fw, float *bw, int capacity, int *steps) {
    return ref_f3d2_run_on((int)PlatformType::Cpu, ref, refData, flo, floData, p, cppOut, fw, bw, capacity, steps);
}

int ref_f3d2_weights(void *hp, float *w) {
    F3d2Handle *h = static_cast<F3d2Handle*>(hp);
    w[0] = h->reg->SimWeight(); w[1] = h->reg->BEWeight(); w[2] = h->reg->LEWeight();
    return 0;
}

} // extern "C"

extern "C" {

// ---------------------------------------------------------------- Aladin (block matching + LTS)

// initialise_block_matching_method + block_matching_method (cpu/_reg_blockMatching.cpp:213-273,
// 455-690) on a reference image and an already-warped floating image.  Outputs are in the
// reference's rank order; counts = {totalBlockNumber, activeBlockNumber, definedActiveBlockNumber}.
int ref_block_matching(const nr_geom *ref, const float *refData, const float *warped, const int *mask,
                       int blockPercent, int inlierLts, int *totalBlock, float *refPos, float *warpPos,
                       int *counts) {
    return guarded([&] {
        NiftiImage r = make_image(ref, refData);
        NiftiImage w = make_image(ref, warped);
        std::vector<int> m(nvox(ref), 0);
        if (mask) std::memcpy(m.data(), mask, m.size() * sizeof(int));
        _reg_blockMatchingParam p;
        initialise_block_matching_method(r, &p, blockPercent, inlierLts, 1, m.data(), false);
        counts[0] = p.totalBlockNumber;
        counts[1] = p.activeBlockNumber;
        if (totalBlock) std::memcpy(totalBlock, p.totalBlock, p.totalBlockNumber * sizeof(int));
        if (warped) {
            block_matching_method(r, w, &p, m.data());
            counts[2] = p.definedActiveBlockNumber;
            if (refPos) std::memcpy(refPos, p.referencePosition, 3 * p.activeBlockNumber * sizeof(float));
            if (warpPos) std::memcpy(warpPos, p.warpedPosition, 3 * p.activeBlockNumber * sizeof(float));
        }
    });
}

// optimize (cpu/_reg_blockMatching.cpp:722-821): LTS update of a 4x4 matrix (row-major, in/out)
// from nActive correspondences of which nDefined have a non-NaN warped position.
int ref_lts(const float *refPos, const float *warpPos, int nActive, int nDefined, int percentToKeep, int affine,
            float *matrix) {
    return guarded([&] {
        _reg_blockMatchingParam p;
        p.dim = 3;
        p.blockNumber[0] = p.blockNumber[1] = p.blockNumber[2] = 2;
        p.activeBlockNumber = nActive;
        p.definedActiveBlockNumber = nDefined;
        p.percent_to_keep = percentToKeep;
        p.referencePosition = (float *)malloc(3 * nActive * sizeof(float));
        p.warpedPosition = (float *)malloc(3 * nActive * sizeof(float));
        std::memcpy(p.referencePosition, refPos, 3 * nActive * sizeof(float));
        std::memcpy(p.warpedPosition, warpPos, 3 * nActive * sizeof(float));
        mat44 m = to_mat44(matrix);
        optimize(&p, &m, affine != 0);
        from_mat44(m, matrix);
    });
}

typedef struct {
    int levels;           // pyramid levels
    int levels_to_perform;
    int max_iterations;   // per level (doubled at the first level)
    int block_percent;
    int inlier_lts;
    int rigid;            // perform the rigid stage
    int affine;           // perform the affine stage
    int align_centre;
    int symmetric;        // reg_aladin_sym (the reg_aladin default) instead of reg_aladin
} ref_aladin_params;

// reg_aladin<float>::Run (reg-lib/_reg_aladin.cpp:315-364) on the given platform; matrix receives
// the final affine.
int ref_aladin_run_on(int platform, const nr_geom *ref, const float *refData, const nr_geom *flo,
                      const float *floData, const ref_aladin_params *p, float *matrix) {
    return guarded([&] {
        std::unique_ptr<reg_aladin<float>> owner(p->symmetric ? new reg_aladin_sym<float>() : new reg_aladin<float>());
        reg_aladin<float>& a = *owner;
        a.SetPlatformType(static_cast<PlatformType>(platform));
        a.SetInputReference(make_image(ref, refData));
        a.SetInputFloating(make_image(flo, floData));
        a.SetNumberOfLevels(p->levels);
        a.SetLevelsToPerform(p->levels_to_perform);
        a.SetMaxIterations(p->max_iterations);
        a.SetBlockPercentage(p->block_percent);
        a.SetInlierLts(p->inlier_lts);
        a.SetPerformRigid(p->rigid != 0);
        a.SetPerformAffine(p->affine != 0);
        a.SetAlignCentre(p->align_centre != 0);
        a.SetVerbose(false);
        a.Run();
        from_mat44(*a.GetTransformationMatrix(), matrix);
    });
}

int ref_aladin_run(const nr_geom *ref, const float *refData, const nr_geom *flo, const float *floData,
                   const ref_aladin_params *p, float *matrix) {
    return ref_aladin_run_on((int)PlatformType::Cpu, ref, refData, flo, floData, p, matrix);
}

// NiftyReg::Mat44Avg2 (reg-lib/cpu/Maths.cpp:479-491): expm((logm(A) + logm(B)) / 2) through Eigen.
int ref_mat44_avg2(const float *A, const float *B, float *out) {
    return guarded([&] {
        mat44 a = to_mat44(A), b = to_mat44(B);
        from_mat44(NiftyReg::Mat44Avg2(&a, &b), out);
    });
}

// nifti_mat44_inverse (reg-io/niftilib/nifti1_io.c:1680-1717).
int ref_mat44_inverse(const float *A, float *out) {
    return guarded([&] { from_mat44(nifti_mat44_inverse(to_mat44(A)), out); });
}

// ---------------------------------------------------------------- file formats
// What the reference's readers see in a NIfTI file: nifti_image_read + reg_checkAndCorrectDimension
// (reg_io_ReadImageFile, reg-io/_reg_ReadWriteImage.cpp:51-81, without its PNG / NRRD branches).
typedef struct {
    int dim[8];
    float pixdim[8];
    int datatype, qform_code, sform_code;
    float quatern[3], qoffset[3], qfac;
    float qto_xyz[16], sto_xyz[16];
    float scl_slope, scl_inter;
    int intent_code;
    float intent_p1;
    char intent_name[16];
} ref_nifti_fields;

static void fields_of(const nifti_image *im, ref_nifti_fields *f) {
    std::memset(f, 0, sizeof(*f));
    for (int i = 0; i < 8; ++i) {
        f->dim[i] = im->dim[i];
        f->pixdim[i] = im->pixdim[i];
    }
    f->datatype = im->datatype;
    f->qform_code = im->qform_code;
    f->sform_code = im->sform_code;
    f->quatern[0] = im->quatern_b; f->quatern[1] = im->quatern_c; f->quatern[2] = im->quatern_d;
    f->qoffset[0] = im->qoffset_x; f->qoffset[1] = im->qoffset_y; f->qoffset[2] = im->qoffset_z;
    f->qfac = im->qfac;
    from_mat44(im->qto_xyz, f->qto_xyz);
    from_mat44(im->sto_xyz, f->sto_xyz);
    f->scl_slope = im->scl_slope;
    f->scl_inter = im->scl_inter;
    f->intent_code = im->intent_code;
    f->intent_p1 = im->intent_p1;
    std::memcpy(f->intent_name, im->intent_name, 16);
}

// Reads a file as the reference does; data (optional, capacity floats) receives the voxels
// converted to float.
int ref_nifti_read(const char *path, ref_nifti_fields *f, float *data, size_t capacity) {
    return guarded([&] {
        nifti_image *im = nifti_image_read(path, data != nullptr);
        if (!im) throw std::runtime_error(std::string("niftilib cannot read ") + path);
        NiftiImage img(im, NiftiImage::Copy::Acquire);
        reg_checkAndCorrectDimension(img);
        fields_of(img, f);
        if (data) {
            reg_tools_changeDatatype<float>(img);
            if (img.nVoxels() > capacity) throw std::runtime_error("capacity");
            std::memcpy(data, img->data, img.nVoxels() * sizeof(float));
        }
    });
}

// reg_createControlPointGrid<float> for the reference image in `refPath` and a spacing in mm
// (as reg_f3d builds its output grid); the grid is written to `gridPath` with niftilib
// (reg_io_WriteImageFile's NIfTI branch) when gridPath is not null.
int ref_control_point_grid(const char *refPath, const float spacing[3], const char *gridPath, ref_nifti_fields *f) {
    return guarded([&] {
        nifti_image *im = nifti_image_read(refPath, 0);
        if (!im) throw std::runtime_error(std::string("niftilib cannot read ") + refPath);
        NiftiImage ref(im, NiftiImage::Copy::Acquire);
        reg_checkAndCorrectDimension(ref);
        NiftiImage grid;
        reg_createControlPointGrid<float>(grid, ref, spacing);
        fields_of(grid, f);
        if (gridPath) {
            nifti_set_filenames(grid, gridPath, 0, 0);
            nifti_image_write(grid);
        }
    });
}

// reg_tool_WriteAffineFile / reg_tool_ReadAffineFile (reg-io/_reg_ReadWriteMatrix.cpp:83-116).
int ref_affine_write(const char *path, const float *m) {
    return guarded([&] {
        const mat44 a = to_mat44(m);
        reg_tool_WriteAffineFile(&a, path);
    });
}

int ref_affine_read(const char *path, float *m) {
    return guarded([&] {
        mat44 a;
        std::string p(path);
        reg_tool_ReadAffineFile(&a, &p[0]);
        from_mat44(a, m);
    });
}

// ---------------------------------------------------------------- LNCC

// reg_lncc value (reg-lib/cpu/_reg_lncc.cpp:303-345) on a reference and a warped image already
// rescaled to [0, 1]; Gaussian kernel, sigma in voxels when negative.
int ref_lncc_value(const nr_geom *ref, const float *refData, const float *warped, const int *mask, float sigma,
                   double *value) {
    return guarded([&] {
        NiftiImage r = make_image(ref, refData), w = make_image(ref, warped);
        NiftiImage mean = make_image(ref, refData), sdev = make_image(ref, refData);
        NiftiImage wmean = make_image(ref, refData), wsdev = make_image(ref, refData), corr = make_image(ref, refData);
        std::vector<int> m(nvox(ref), 0), fw(nvox(ref), 0);
        if (mask) std::memcpy(m.data(), mask, m.size() * sizeof(int));
        const float sd[1] = { sigma };
        const double tw[1] = { 1.0 };
        *value = GetSimilarityMeasureValue(r, mean, sdev, w, wmean, wsdev, m.data(), fw.data(), sd, corr,
                                           ConvKernelType::Gaussian, 1, tw);
    });
}

// reg_lncc voxel-based gradient (reg-lib/cpu/_reg_lncc.cpp:514-577): grad (3N, in/out) -= ...
int ref_lncc_gradient(const nr_geom *ref, const float *refData, const float *warped, const float *warpedGrad,
                      const int *mask, float sigma, double weight, float *grad) {
    return guarded([&] {
        NiftiImage r = make_image(ref, refData), w = make_image(ref, warped);
        NiftiImage mean = make_image(ref, refData), sdev = make_image(ref, refData);
        NiftiImage wmean = make_image(ref, refData), wsdev = make_image(ref, refData), corr = make_image(ref, refData);
        NiftiImage wg = make_vector_image(ref, warpedGrad), g = make_vector_image(ref, grad);
        std::vector<int> m(nvox(ref), 0), fw(nvox(ref), 0);
        if (mask) std::memcpy(m.data(), mask, m.size() * sizeof(int));
        const float sd[1] = { sigma };
        GetVoxelBasedSimilarityMeasureGradient(r, mean, sdev, w, wmean, wsdev, m.data(), fw.data(), sd, corr, wg, g,
                                               ConvKernelType::Gaussian, 0, weight);
        std::memcpy(grad, g->data, 3 * nvox(ref) * sizeof(float));
    });
}

} // extern "C"
