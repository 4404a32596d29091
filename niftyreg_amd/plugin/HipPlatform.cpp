// HIP platform classes over libnrhip's C ABI; see HipPlatform.hpp.
#include "HipPlatform.hpp"

#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>

namespace NiftyRegHip {

void Check(int rc, const char *call) {
    if (rc != NRH_OK) NR_FATAL_ERROR(std::string(call) + ": " + nrh_last_error());
}

nrh_geom GeomOf(const nifti_image *img) {
    nrh_geom g{};
    g.nx = img->nx;
    g.ny = img->ny;
    g.nz = img->nz;
    g.dx = img->dx;
    g.dy = img->dy;
    g.dz = img->dz;
    const bool s = img->sform_code > 0;
    const mat44& xyz = s ? img->sto_xyz : img->qto_xyz;
    const mat44& ijk = s ? img->sto_ijk : img->qto_ijk;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            g.xyz[r * 4 + c] = xyz.m[r][c];
            g.ijk[r * 4 + c] = ijk.m[r][c];
        }
    g.sform = s ? 1 : 0;
    return g;
}

static void ToRowMajor(const mat44& m, float out[16]) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[r * 4 + c] = m.m[r][c];
}

}  // namespace NiftyRegHip

using NiftyRegHip::Check;
using NiftyRegHip::GeomOf;

// ---------------------------------------------------------------- HipF3dContent
HipF3dContent::HipF3dContent(NiftiImage& referenceIn, NiftiImage& floatingIn, NiftiImage& controlPointGridIn,
                             NiftiImage *localWeightSimIn, int *referenceMaskIn, mat44 *transformationMatrixIn,
                             size_t bytesIn):
    F3dContent(referenceIn, floatingIn, controlPointGridIn, localWeightSimIn, referenceMaskIn, transformationMatrixIn,
               sizeof(float)),
    DefContent(referenceIn, floatingIn, localWeightSimIn, referenceMaskIn, transformationMatrixIn, sizeof(float)),
    Content(referenceIn, floatingIn, referenceMaskIn, transformationMatrixIn, sizeof(float)) {
    if (bytesIn != sizeof(float)) NR_FATAL_ERROR("The HIP platform computes in single precision");
    // reg_base passes its (usually empty) weight image; only a set image is a local weighting
    if (localWeightSimIn && *localWeightSimIn)
        NR_FATAL_ERROR("Local similarity weights are not supported on the HIP platform");
    params.ref_bins = params.flo_bins = 68;
    params.padding = std::numeric_limits<float>::quiet_NaN();
    params.similarity_weight = 1.f;
    params.be_weight = 0.f;
    params.le_weight = 0.f;
    params.image_storage = 0;
    params.measure = NRH_MEASURE_NMI;
    params.lncc_sigma = -5.f;
    // the field rounding of the build this plugin is compiled into
#ifdef USE_SSE
    params.field_order = NRH_FIELD_ORDER_SSE;
#else
    params.field_order = NRH_FIELD_ORDER_SCALAR;
#endif
    const nrh_geom r = GeomOf(reference), f = GeomOf(floating), c = GeomOf(controlPointGrid);
    Check(nrh_f3d_create(&r, static_cast<const float*>(reference->data), &f, static_cast<const float*>(floating->data),
                         referenceMask, &c, &params, &handle), "nrh_f3d_create");
    const char *fused = std::getenv("NRH_PLUGIN_FUSED");
    fusing = fused && std::strcmp(fused, "1") == 0;
    HipF3dContent::UpdateControlPointGrid();
}

void HipF3dContent::Materialise() {
    const int p = pending;
    if (!p) return;
    pending = 0;
    if (p & SField) {
        Check(nrh_f3d_deformation_field(handle), "nrh_f3d_deformation_field");
        onDevice |= Field;
    }
    if (p & SWarp) {
        Check(nrh_f3d_resample(handle), "nrh_f3d_resample");
        onDevice |= Warped;
    }
    if (p & SImageGradient) {
        Check(nrh_f3d_image_gradient(handle), "nrh_f3d_image_gradient");
        onDevice |= WarpedGradient;
    }
    if (p & SZeroVoxel) {
        Check(nrh_f3d_zero_voxel_gradient(handle), "nrh_f3d_zero_voxel_gradient");
        onDevice |= VoxelGradient;
    }
    if (p & SVoxelGradient) {
        // the measure's voxel gradient at time-point weight 1 (the only weight the route defers)
        if (params.measure == NRH_MEASURE_LNCC) {
            Check(nrh_f3d_lncc_voxel_gradient(handle, 1.0), "nrh_f3d_lncc_voxel_gradient");
        } else {
            double nmi = 0, ent[4];
            Check(nrh_f3d_nmi_value(handle, &nmi, ent), "nrh_f3d_nmi_value");
            Check(nrh_f3d_nmi_voxel_gradient(handle, 1.0), "nrh_f3d_nmi_voxel_gradient");
        }
    }
    if (p & SConvolved) Check(nrh_f3d_convolve_voxel_gradient(handle), "nrh_f3d_convolve_voxel_gradient");
}

HipF3dContent::~HipF3dContent() {
    if (handle) {
        HipF3dContent::GetControlPointGrid();  // leave the final positions in the host image
        nrh_f3d_destroy(handle);
    }
}

void HipF3dContent::SetMeasure(int measure, int refBins, int floBins, float lnccSigma) {
    params.measure = measure;
    params.ref_bins = refBins;
    params.flo_bins = floBins;
    params.lncc_sigma = lnccSigma;
    Check(nrh_f3d_set_params(handle, &params), "nrh_f3d_set_params");
}

NiftiImage& HipF3dContent::GetControlPointGrid() {
    Check(nrh_f3d_get_control_points(handle, static_cast<float*>(controlPointGrid->data)), "nrh_f3d_get_control_points");
    return controlPointGrid;
}

NiftiImage& HipF3dContent::GetTransformationGradient() {
    Check(nrh_f3d_download_node_gradient(handle, static_cast<float*>(transformationGradient->data)),
          "nrh_f3d_download_node_gradient");
    return transformationGradient;
}

NiftiImage& HipF3dContent::GetDeformationField() {
    Materialise();
    if (onDevice & Field)
        Check(nrh_f3d_download_field(handle, static_cast<float*>(deformationField->data)), "nrh_f3d_download_field");
    return deformationField;
}

NiftiImage& HipF3dContent::GetWarped() {
    Materialise();
    if (onDevice & Warped)
        Check(nrh_f3d_download_warped(handle, static_cast<float*>(warped->data)), "nrh_f3d_download_warped");
    return warped;
}

NiftiImage& HipF3dContent::GetWarpedGradient() {
    Materialise();
    if (onDevice & WarpedGradient)
        Check(nrh_f3d_download_warped_gradient(handle, static_cast<float*>(warpedGradient->data)),
          "nrh_f3d_download_warped_gradient");
    return warpedGradient;
}

NiftiImage& HipF3dContent::GetVoxelBasedMeasureGradient() {
    Materialise();
    if (onDevice & VoxelGradient)
        Check(nrh_f3d_download_voxel_gradient(handle, static_cast<float*>(voxelBasedMeasureGradient->data)),
          "nrh_f3d_download_voxel_gradient");
    return voxelBasedMeasureGradient;
}

void HipF3dContent::UpdateControlPointGrid() {
    Check(nrh_f3d_set_control_points(handle, static_cast<const float*>(controlPointGrid->data)),
          "nrh_f3d_set_control_points");
}

void HipF3dContent::UpdateTransformationGradient() {
    Check(nrh_f3d_upload_node_gradient(handle, static_cast<const float*>(transformationGradient->data)),
          "nrh_f3d_upload_node_gradient");
}

void HipF3dContent::ZeroTransformationGradient() {
    Check(nrh_f3d_zero_node_gradient(handle), "nrh_f3d_zero_node_gradient");
}

void HipF3dContent::ZeroVoxelBasedMeasureGradient() {
    if ((pending & (SField | SWarp)) == (SField | SWarp) && !(pending & ~(SField | SWarp | SImageGradient))) {
        Defer(SZeroVoxel);
        return;
    }
    Materialise();
    Check(nrh_f3d_zero_voxel_gradient(handle), "nrh_f3d_zero_voxel_gradient");
    onDevice |= VoxelGradient;
}

void HipF3dContent::SetDeformationField(NiftiImage&& deformationFieldIn) {
    Materialise();
    F3dContent::SetDeformationField(std::move(deformationFieldIn));
    onDevice &= ~Field;
    EnsureFieldOnDevice();
}

void HipF3dContent::UpdateDeformationField() {
    Materialise();
    onDevice &= ~Field;
    EnsureFieldOnDevice();
}

void HipF3dContent::UpdateVoxelBasedMeasureGradient() {
    Materialise();
    onDevice &= ~VoxelGradient;
    EnsureVoxelGradientOnDevice();
}

void HipF3dContent::EnsureFieldOnDevice() {
    Materialise();
    if (onDevice & Field) return;
    Check(nrh_f3d_upload_field(handle, static_cast<const float*>(deformationField->data)), "nrh_f3d_upload_field");
    onDevice |= Field;
}

void HipF3dContent::EnsureVoxelGradientOnDevice() {
    Materialise();
    if (onDevice & VoxelGradient) return;
    Check(nrh_f3d_upload_voxel_gradient(handle, static_cast<const float*>(voxelBasedMeasureGradient->data)),
          "nrh_f3d_upload_voxel_gradient");
    onDevice |= VoxelGradient;
}

void HipF3dContent::SyncVelocityGrid() {
    if (controlPointGrid->intent_p1 != SPLINE_VEL_GRID) NR_FATAL_ERROR("The control point grid is not a velocity grid");
    Check(nrh_f3d_set_velocity_grid(handle, (int)controlPointGrid->intent_p2), "nrh_f3d_set_velocity_grid");
    // the grid's extensions 0 / 1 (half affines) and the content's transformation matrix
    float ext[2][16], trans[16];
    const float *e[2] = { nullptr, nullptr };
    for (int k = 0; k < 2 && k < controlPointGrid->num_ext; ++k) {
        if (!controlPointGrid->ext_list[k].edata) continue;
        const mat44 *m = reinterpret_cast<const mat44*>(controlPointGrid->ext_list[k].edata);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) ext[k][i * 4 + j] = m->m[i][j];
        e[k] = ext[k];
    }
    if (transformationMatrix)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) trans[i * 4 + j] = transformationMatrix->m[i][j];
    Check(nrh_f3d_set_velocity_affines(handle, e[0], e[1], transformationMatrix ? trans : nullptr),
          "nrh_f3d_set_velocity_affines");
}

// ---------------------------------------------------------------- HipAladinContent
HipAladinContent::HipAladinContent(NiftiImage& referenceIn, NiftiImage& floatingIn, int *referenceMaskIn,
                                   mat44 *transformationMatrixIn, size_t bytesIn, const unsigned percentageOfBlocks,
                                   const unsigned inlierLtsIn, int blockStepSize):
    AladinContent(referenceIn, floatingIn, referenceMaskIn, transformationMatrixIn, sizeof(float), percentageOfBlocks,
                  inlierLtsIn, blockStepSize),
    Content(referenceIn, floatingIn, referenceMaskIn, transformationMatrixIn, sizeof(float)) {
    if (bytesIn != sizeof(float)) NR_FATAL_ERROR("The HIP platform computes in single precision");
    if (!blockMatchingParams) return;  // the final-warp content of reg_aladin carries no blocks
    if (blockStepSize != 1) NR_FATAL_ERROR("The HIP block matching uses a step of 1 voxel");
    const nrh_aladin_params p{ (int)percentageOfBlocks, (int)inlierLtsIn, std::numeric_limits<float>::quiet_NaN() };
    const nrh_geom r = GeomOf(reference), f = GeomOf(floating);
    Check(nrh_aladin_create(&r, static_cast<const float*>(reference->data), &f,
                            static_cast<const float*>(floating->data), referenceMask, &p, &handle),
          "nrh_aladin_create");
}

HipAladinContent::~HipAladinContent() {
    if (handle) nrh_aladin_destroy(handle);
}

NiftiImage& HipAladinContent::GetWarped() {
    if (handle) Check(nrh_aladin_download_warped(handle, static_cast<float*>(warped->data)), "nrh_aladin_download_warped");
    return warped;
}

_reg_blockMatchingParam* HipAladinContent::GetBlockMatchingParams() {
    if (handle && blockMatchingParams) {
        int defined = 0;
        Check(nrh_aladin_download_positions(handle, blockMatchingParams->referencePosition,
                                            blockMatchingParams->warpedPosition, &defined),
              "nrh_aladin_download_positions");
        blockMatchingParams->definedActiveBlockNumber = defined;
    }
    return blockMatchingParams;
}

ContentCreator* HipContentCreatorFactory::Produce(const ContentType conType) {
    switch (conType) {
    case ContentType::F3d:
        return new HipF3dContentCreator();
    case ContentType::F3d2:
        return new HipF3d2ContentCreator();
    case ContentType::Aladin:
        return new HipAladinContentCreator();
    default:
        NR_FATAL_ERROR("This content type is not supported on the HIP platform");
        return nullptr;
    }
}

// ---------------------------------------------------------------- HipCompute
nrh_f3d* HipCompute::Handle() {
    return dynamic_cast<HipF3dContent&>(con).GetHandle();
}

void HipCompute::ResampleImage(int interpolation, float paddingValue) {
    if (interpolation != 1) NR_FATAL_ERROR("Only linear interpolation is supported on the HIP platform");
    if (!(paddingValue != paddingValue)) NR_FATAL_ERROR("The HIP resampling pads with NaN");
    HipF3dContent& c = dynamic_cast<HipF3dContent&>(con);
    if (c.Pending() == HipF3dContent::SField) {
        c.Defer(HipF3dContent::SWarp);
        return;
    }
    c.Materialise();
    Check(nrh_f3d_resample(Handle()), "nrh_f3d_resample");
    c.MarkOnDevice(HipF3dContent::Warped);
}

double HipCompute::ApproxBendingEnergy() {
    double v = 0;
    Check(nrh_f3d_bending_energy(Handle(), &v), "nrh_f3d_bending_energy");
    return v;
}

void HipCompute::ApproxBendingEnergyGradient(float weight) {
    Check(nrh_f3d_bending_energy_gradient(Handle(), weight), "nrh_f3d_bending_energy_gradient");
}

double HipCompute::ApproxLinearEnergy() {
    double v = 0;
    Check(nrh_f3d_linear_energy(Handle(), &v), "nrh_f3d_linear_energy");
    return v;
}

void HipCompute::ApproxLinearEnergyGradient(float weight) {
    Check(nrh_f3d_linear_energy_gradient(Handle(), weight), "nrh_f3d_linear_energy_gradient");
}

void HipCompute::GetDeformationField(bool composition, bool bspline) {
    if (composition || !bspline) NR_FATAL_ERROR("Only the non-composed cubic B-spline field is supported on HIP");
    HipF3dContent& c = dynamic_cast<HipF3dContent&>(con);
    if (c.Fusing()) {
        // the stages recorded so far belong to the same control points and were not read
        c.DropPending();
        c.Defer(HipF3dContent::SField);
        return;
    }
    c.Materialise();
    Check(nrh_f3d_deformation_field(Handle()), "nrh_f3d_deformation_field");
    c.MarkOnDevice(HipF3dContent::Field);
}

void HipCompute::UpdateControlPointPosition(float*, const float*, const float*, const float scale,
                                           const bool optimiseX, const bool optimiseY, const bool optimiseZ) {
    if (!optimiseX || !optimiseY || !optimiseZ) NR_FATAL_ERROR("The HIP optimiser updates every axis");
    Check(nrh_f3d_update_control_points(Handle(), scale), "nrh_f3d_update_control_points");
}

void HipCompute::GetImageGradient(int interpolation, float, int activeTimePoint) {
    if (interpolation != 1 || activeTimePoint != 0) NR_FATAL_ERROR("Linear, single time point gradient only on HIP");
    HipF3dContent& c = dynamic_cast<HipF3dContent&>(con);
    const int fw = HipF3dContent::SField | HipF3dContent::SWarp;
    if (c.Pending() == fw || c.Pending() == (fw | HipF3dContent::SZeroVoxel)) {
        c.Defer(HipF3dContent::SImageGradient);
        return;
    }
    c.Materialise();
    Check(nrh_f3d_image_gradient(Handle()), "nrh_f3d_image_gradient");
    c.MarkOnDevice(HipF3dContent::WarpedGradient);
}

double HipCompute::GetMaximalLength(bool optimiseX, bool optimiseY, bool optimiseZ) {
    if (!optimiseX || !optimiseY || !optimiseZ) NR_FATAL_ERROR("The HIP optimiser updates every axis");
    double v = 0;
    Check(nrh_f3d_max_length(Handle(), &v), "nrh_f3d_max_length");
    return v;
}

void HipCompute::NormaliseGradient(double maxGradLength, bool optimiseX, bool optimiseY, bool optimiseZ) {
    if (!optimiseX || !optimiseY || !optimiseZ) NR_FATAL_ERROR("The HIP optimiser updates every axis");
    Check(nrh_f3d_normalise_gradient(Handle(), maxGradLength), "nrh_f3d_normalise_gradient");
}

void HipCompute::ConvolveVoxelBasedMeasureGradient(float weight) {
    HipF3dContent& c = dynamic_cast<HipF3dContent&>(con);
    const int all = HipF3dContent::SField | HipF3dContent::SWarp | HipF3dContent::SImageGradient |
                    HipF3dContent::SZeroVoxel | HipF3dContent::SVoxelGradient;
    if (c.Pending() == all) {
        // the whole similarity gradient in one call; the voxel gradient stays recorded (convolved)
        Check(nrh_f3d_similarity_gradient(Handle(), weight), "nrh_f3d_similarity_gradient");
        c.Defer(HipF3dContent::SConvolved);
        return;
    }
    c.Materialise();
    Check(nrh_f3d_convolve_voxel_gradient(Handle()), "nrh_f3d_convolve_voxel_gradient");
    VoxelCentricToNodeCentric(weight);
}

void HipCompute::VoxelCentricToNodeCentric(float weight) {
    dynamic_cast<HipF3dContent&>(con).Materialise();
    Check(nrh_f3d_voxel_to_node(Handle(), weight), "nrh_f3d_voxel_to_node");
}

void HipCompute::GetAffineDeformationField(bool) {
    NR_FATAL_ERROR("Affine fields of an F3D content are not computed on the HIP platform");
}

// Terms reg_f3d only calls when their weights are non-zero (-jl, -land, -smoothGrad, F3D2): the
// device holds the current control points, so the base class's CPU code would read stale host
// data; they fail instead.
double HipCompute::GetJacobianPenaltyTerm(bool) { NR_FATAL_ERROR("The Jacobian penalty is not implemented on HIP"); return 0; }
void HipCompute::JacobianPenaltyTermGradient(float, bool) { NR_FATAL_ERROR("The Jacobian penalty is not implemented on HIP"); }
double HipCompute::CorrectFolding(bool) { NR_FATAL_ERROR("Folding correction is not implemented on HIP"); return 0; }
double HipCompute::GetLandmarkDistance(size_t, float *, float *) {
    NR_FATAL_ERROR("Landmark terms are not implemented on HIP");
    return 0;
}
void HipCompute::LandmarkDistanceGradient(size_t, float *, float *, float) {
    NR_FATAL_ERROR("Landmark terms are not implemented on HIP");
}
void HipCompute::SmoothGradient(float) { NR_FATAL_ERROR("Gradient smoothing is not implemented on HIP"); }
void HipCompute::GetApproximatedGradient(InterfaceOptimiser&) {
    NR_FATAL_ERROR("The finite-difference gradient is not implemented on HIP");
}
// ---- velocity grids (reg_f3d2)
void HipCompute::GetDefFieldFromVelocityGrid(const bool updateStepNumber) {
    HipF3dContent& c = dynamic_cast<HipF3dContent&>(con);
    if (c.GridHeader()->intent_p1 == CUB_SPLINE_GRID) {
        // reg_spline_getDefFieldFromVelocityGrid on a displacement grid: the plain spline field
        GetDeformationField(false, true);
        return;
    }
    c.SyncVelocityGrid();
    Check(nrh_f3d_def_field_from_velocity(Handle(), updateStepNumber ? 1 : 0), "nrh_f3d_def_field_from_velocity");
    int steps = 0;
    Check(nrh_f3d_get_velocity_steps(Handle(), &steps), "nrh_f3d_get_velocity_steps");
    c.GridHeader()->intent_p2 = (float)steps;
    c.MarkOnDevice(HipF3dContent::Field);
}

void HipCompute::ExponentiateGradient(Content& conBwIn) {
    HipF3dContent& c = dynamic_cast<HipF3dContent&>(con);
    HipF3dContent& b = dynamic_cast<HipF3dContent&>(conBwIn);
    b.SyncVelocityGrid();
    c.EnsureVoxelGradientOnDevice();
    Check(nrh_f3d_exponentiate_gradient(Handle(), b.GetHandle()), "nrh_f3d_exponentiate_gradient");
    c.MarkOnDevice(HipF3dContent::VoxelGradient);
}

void HipCompute::UpdateVelocityField(float scale, bool optimiseX, bool optimiseY, bool optimiseZ) {
    dynamic_cast<HipF3dContent&>(con).SyncVelocityGrid();
    Check(nrh_f3d_update_velocity_field(Handle(), scale, optimiseX, optimiseY, optimiseZ),
          "nrh_f3d_update_velocity_field");
}

void HipCompute::BchUpdate(float scale, int bchUpdateValue) {
    Check(nrh_f3d_bch_update(Handle(), scale, bchUpdateValue), "nrh_f3d_bch_update");
}

void HipCompute::SymmetriseVelocityFields(Content& conBwIn) {
    HipF3dContent& b = dynamic_cast<HipF3dContent&>(conBwIn);
    dynamic_cast<HipF3dContent&>(con).SyncVelocityGrid();
    b.SyncVelocityGrid();
    Check(nrh_f3d_symmetrise_velocity_fields(Handle(), b.GetHandle()), "nrh_f3d_symmetrise_velocity_fields");
}

void HipCompute::DefFieldCompose(const NiftiImage& defField) {
    HipF3dContent& c = dynamic_cast<HipF3dContent&>(con);
    c.EnsureFieldOnDevice();
    const nrh_geom g = GeomOf(defField);
    Check(nrh_f3d_compose_field(Handle(), &g, static_cast<const float*>(defField->data)), "nrh_f3d_compose_field");
    c.MarkOnDevice(HipF3dContent::Field);
}

NiftiImage HipCompute::ResampleGradient(int interpolation, float padding) {
    HipF3dContent& c = dynamic_cast<HipF3dContent&>(con);
    c.EnsureFieldOnDevice();
    c.EnsureVoxelGradientOnDevice();
    NiftiImage warped(c.DefContent::GetVoxelBasedMeasureGradient(), NiftiImage::Copy::ImageInfoAndAllocData);
    Check(nrh_f3d_resample_gradient(Handle(), interpolation, padding, static_cast<float*>(warped->data)),
          "nrh_f3d_resample_gradient");
    return warped;
}

// ---------------------------------------------------------------- measures
void reg_nmi_hip::InitialiseMeasure(nifti_image *refImg, nifti_image *floImg, int *refMask, nifti_image *warpedImg,
                                    nifti_image *warpedGrad, nifti_image *voxelBasedGrad, nifti_image *localWeightSim,
                                    int *floMask, nifti_image *warpedImgBw, nifti_image *warpedGradBw,
                                    nifti_image *voxelBasedGradBw) {
    reg_nmi::InitialiseMeasure(refImg, floImg, refMask, warpedImg, warpedGrad, voxelBasedGrad, localWeightSim, floMask,
                               warpedImgBw, warpedGradBw, voxelBasedGradBw);
    if (!content) NR_FATAL_ERROR("reg_nmi_hip needs a HIP F3D content");
    if (referenceTimePoints != 1) NR_FATAL_ERROR("HIP NMI: one time point");
    if (isSymmetric && !contentBw) NR_FATAL_ERROR("HIP NMI: the symmetric measure needs the backward HIP content");
    const int rb = referenceBinNumber[0], fb = floatingBinNumber[0];
    content->SetMeasure(NRH_MEASURE_NMI, rb, fb, -5.f);
    // the host images were rescaled to [2, bins - 3] by reg_nmi; the device copies follow
    Check(nrh_f3d_rescale_intensities(content->GetHandle(), 2.f, (float)rb - 3.f, 2.f, (float)fb - 3.f),
          "nrh_f3d_rescale_intensities");
    if (isSymmetric) {
        // the backward content's reference is the floating image (reg-lib/cpu/_reg_nmi.cpp:400-413)
        contentBw->SetMeasure(NRH_MEASURE_NMI, fb, rb, -5.f);
        Check(nrh_f3d_rescale_intensities(contentBw->GetHandle(), 2.f, (float)fb - 3.f, 2.f, (float)rb - 3.f),
              "nrh_f3d_rescale_intensities");
    }
}

double reg_nmi_hip::GetSimilarityMeasureValueFw() {
    double nmi = 0, ent[4];
    if (!isSymmetric && content->Pending() == (HipF3dContent::SField | HipF3dContent::SWarp)) {
        Check(nrh_f3d_similarity_value(content->GetHandle(), &nmi), "nrh_f3d_similarity_value");
        return timePointWeights[0] * nmi;
    }
    content->Materialise();
    Check(nrh_f3d_nmi_value(content->GetHandle(), &nmi, ent), "nrh_f3d_nmi_value");
    return timePointWeights[0] * nmi;
}

double reg_nmi_hip::GetSimilarityMeasureValueBw() {
    double nmi = 0, ent[4];
    contentBw->Materialise();
    Check(nrh_f3d_nmi_value(contentBw->GetHandle(), &nmi, ent), "nrh_f3d_nmi_value");
    return timePointWeights[0] * nmi;
}

void reg_nmi_hip::GetVoxelBasedSimilarityMeasureGradientFw(int currentTimePoint) {
    // reg_nmi::GetVoxelBasedSimilarityMeasureGradientFw (reg-lib/cpu/_reg_nmi.cpp:622-637): the
    // joint histograms (both directions when symmetric) of the current warped images first
    const int staged = HipF3dContent::SField | HipF3dContent::SWarp | HipF3dContent::SImageGradient |
                       HipF3dContent::SZeroVoxel;
    if (!isSymmetric && timePointWeights[currentTimePoint] == 1 && content->Pending() == staged) {
        content->Defer(HipF3dContent::SVoxelGradient);
        return;
    }
    this->GetSimilarityMeasureValue();
    content->EnsureVoxelGradientOnDevice();
    Check(nrh_f3d_nmi_voxel_gradient(content->GetHandle(), timePointWeights[currentTimePoint]),
          "nrh_f3d_nmi_voxel_gradient");
    content->MarkOnDevice(HipF3dContent::VoxelGradient);
}

void reg_nmi_hip::GetVoxelBasedSimilarityMeasureGradientBw(int currentTimePoint) {
    contentBw->EnsureVoxelGradientOnDevice();
    Check(nrh_f3d_nmi_voxel_gradient(contentBw->GetHandle(), timePointWeights[currentTimePoint]),
          "nrh_f3d_nmi_voxel_gradient");
    contentBw->MarkOnDevice(HipF3dContent::VoxelGradient);
}

void reg_lncc_hip::InitialiseMeasure(nifti_image *refImg, nifti_image *floImg, int *refMask, nifti_image *warpedImg,
                                     nifti_image *warpedGrad, nifti_image *voxelBasedGrad, nifti_image *localWeightSim,
                                     int *floMask, nifti_image *warpedImgBw, nifti_image *warpedGradBw,
                                     nifti_image *voxelBasedGradBw) {
    reg_lncc::InitialiseMeasure(refImg, floImg, refMask, warpedImg, warpedGrad, voxelBasedGrad, localWeightSim,
                                floMask, warpedImgBw, warpedGradBw, voxelBasedGradBw);
    if (!content) NR_FATAL_ERROR("reg_lncc_hip needs a HIP F3D content");
    if (referenceTimePoints != 1) NR_FATAL_ERROR("HIP LNCC: one time point");
    if (isSymmetric && !contentBw) NR_FATAL_ERROR("HIP LNCC: the symmetric measure needs the backward HIP content");
    if (kernelType != ConvKernelType::Gaussian) NR_FATAL_ERROR("HIP LNCC uses the Gaussian kernel");
    content->SetMeasure(NRH_MEASURE_LNCC, 68, 68, kernelStandardDeviation[0]);
    // reg_lncc rescaled the host images to [0, 1]; the device copies follow
    Check(nrh_f3d_rescale_intensities(content->GetHandle(), 0.f, 1.f, 0.f, 1.f), "nrh_f3d_rescale_intensities");
    if (isSymmetric) {
        // the backward content's reference is the floating image (reg-lib/cpu/_reg_lncc.cpp:376-390)
        contentBw->SetMeasure(NRH_MEASURE_LNCC, 68, 68, kernelStandardDeviation[0]);
        Check(nrh_f3d_rescale_intensities(contentBw->GetHandle(), 0.f, 1.f, 0.f, 1.f), "nrh_f3d_rescale_intensities");
    }
}

double reg_lncc_hip::GetSimilarityMeasureValueFw() {
    double v = 0;
    if (!isSymmetric && content->Pending() == (HipF3dContent::SField | HipF3dContent::SWarp)) {
        Check(nrh_f3d_similarity_value(content->GetHandle(), &v), "nrh_f3d_similarity_value");
        return timePointWeights[0] * v;
    }
    content->Materialise();
    Check(nrh_f3d_lncc_value(content->GetHandle(), &v), "nrh_f3d_lncc_value");
    return timePointWeights[0] * v;
}

double reg_lncc_hip::GetSimilarityMeasureValueBw() {
    double v = 0;
    contentBw->Materialise();
    Check(nrh_f3d_lncc_value(contentBw->GetHandle(), &v), "nrh_f3d_lncc_value");
    return timePointWeights[0] * v;
}

void reg_lncc_hip::GetVoxelBasedSimilarityMeasureGradientFw(int currentTimePoint) {
    const int staged = HipF3dContent::SField | HipF3dContent::SWarp | HipF3dContent::SImageGradient |
                       HipF3dContent::SZeroVoxel;
    if (!isSymmetric && timePointWeights[currentTimePoint] == 1 && content->Pending() == staged) {
        content->Defer(HipF3dContent::SVoxelGradient);
        return;
    }
    content->EnsureVoxelGradientOnDevice();
    Check(nrh_f3d_lncc_voxel_gradient(content->GetHandle(), timePointWeights[currentTimePoint]),
          "nrh_f3d_lncc_voxel_gradient");
    content->MarkOnDevice(HipF3dContent::VoxelGradient);
}

void reg_lncc_hip::GetVoxelBasedSimilarityMeasureGradientBw(int currentTimePoint) {
    contentBw->EnsureVoxelGradientOnDevice();
    Check(nrh_f3d_lncc_voxel_gradient(contentBw->GetHandle(), timePointWeights[currentTimePoint]),
          "nrh_f3d_lncc_voxel_gradient");
    contentBw->MarkOnDevice(HipF3dContent::VoxelGradient);
}

reg_measure* HipMeasureCreator::Create(const MeasureType measureType) {
    switch (measureType) {
    case MeasureType::Nmi:
        return new reg_nmi_hip();
    case MeasureType::Lncc:
        return new reg_lncc_hip();
    default:
        NR_FATAL_ERROR("Only NMI and LNCC are implemented on the HIP platform");
        return nullptr;
    }
}

void HipMeasureCreator::Initialise(reg_measure& measure, DefContent& con, DefContent *conBw) {
    HipF3dContent *c = dynamic_cast<HipF3dContent*>(&con);
    HipF3dContent *cBw = conBw ? dynamic_cast<HipF3dContent*>(conBw) : nullptr;
    if (conBw && !cBw) NR_FATAL_ERROR("The backward content is not a HIP F3D content");
    if (auto *m = dynamic_cast<reg_nmi_hip*>(&measure)) m->SetContent(c, cBw);
    if (auto *m = dynamic_cast<reg_lncc_hip*>(&measure)) m->SetContent(c, cBw);
    MeasureCreator::Initialise(measure, con, conBw);
}

// ---------------------------------------------------------------- Aladin kernels
static HipAladinContent* AladinOf(Content *con) {
    auto *c = dynamic_cast<HipAladinContent*>(con);
    if (!c || !c->GetHandle()) NR_FATAL_ERROR("The HIP Aladin kernels need a HIP Aladin content with blocks");
    return c;
}

HipAffineDeformationFieldKernel::HipAffineDeformationFieldKernel(Content *con): content(AladinOf(con)) {}

void HipAffineDeformationFieldKernel::Calculate(bool compose) {
    if (compose) NR_FATAL_ERROR("Composed affine fields are not supported on the HIP platform");
    // the field is evaluated inside the warp; only the matrix moves to the device here
    float m[16];
    NiftyRegHip::ToRowMajor(*content->GetTransformationMatrix(), m);
    Check(nrh_aladin_set_matrix(content->GetHandle(), m), "nrh_aladin_set_matrix");
}

HipResampleImageKernel::HipResampleImageKernel(Content *con): content(AladinOf(con)) {}

void HipResampleImageKernel::Calculate(int interp, float paddingValue, bool *, mat33 *) {
    if (interp != 1) NR_FATAL_ERROR("Only linear interpolation is supported on the HIP platform");
    if (!(paddingValue != paddingValue)) NR_FATAL_ERROR("The HIP resampling pads with NaN");
    Check(nrh_aladin_warp(content->GetHandle()), "nrh_aladin_warp");
}

HipBlockMatchingKernel::HipBlockMatchingKernel(Content *con): content(AladinOf(con)) {}

void HipBlockMatchingKernel::Calculate() {
    Check(nrh_aladin_block_matching(content->GetHandle()), "nrh_aladin_block_matching");
}

HipLtsKernel::HipLtsKernel(Content *con): content(AladinOf(con)) {}

void HipLtsKernel::Calculate(bool affine) {
    Check(nrh_aladin_lts(content->GetHandle(), affine ? 1 : 0), "nrh_aladin_lts");
    float m[16];
    Check(nrh_aladin_get_matrix(content->GetHandle(), m), "nrh_aladin_get_matrix");
    mat44 *t = content->GetTransformationMatrix();
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) t->m[r][c] = m[r * 4 + c];
}

void HipConvolutionKernel::Calculate(nifti_image*, float*, ConvKernelType, int*, bool*, bool*) {
    NR_FATAL_ERROR("Input smoothing (-smooR / -smooF) is not implemented on the HIP platform");
}

Kernel* HipKernelFactory::Produce(std::string name, Content *con) const {
    if (name == AffineDeformationFieldKernel::GetName()) return new HipAffineDeformationFieldKernel(con);
    if (name == ConvolutionKernel::GetName()) return new HipConvolutionKernel();
    if (name == BlockMatchingKernel::GetName()) return new HipBlockMatchingKernel(con);
    if (name == ResampleImageKernel::GetName()) return new HipResampleImageKernel(con);
    if (name == LtsKernel::GetName()) return new HipLtsKernel(con);
    return nullptr;
}

// ---------------------------------------------------------------- optimiser
void HipConjugateGradient::StoreCurrentDof() {
    Check(nrh_f3d_store_best(content.GetHandle()), "nrh_f3d_store_best");
    if (contentBw) Check(nrh_f3d_store_best(contentBw->GetHandle()), "nrh_f3d_store_best");
}

void HipConjugateGradient::RestoreBestDof() {
    Check(nrh_f3d_restore_best(content.GetHandle()), "nrh_f3d_restore_best");
    if (contentBw) Check(nrh_f3d_restore_best(contentBw->GetHandle()), "nrh_f3d_restore_best");
}

void HipConjugateGradient::UpdateGradientValues() {
    if (contentBw)
        Check(nrh_f3d2_conjugate_gradient(content.GetHandle(), contentBw->GetHandle(), this->firstCall ? 1 : 0),
              "nrh_f3d2_conjugate_gradient");
    else
        Check(nrh_f3d_conjugate_gradient(content.GetHandle(), this->firstCall ? 1 : 0), "nrh_f3d_conjugate_gradient");
    this->firstCall = false;
}

void HipConjugateGradient::Perturbation(float length) {
    content.GetControlPointGrid();  // the host arrays the optimiser was initialised with
    if (contentBw) contentBw->GetControlPointGrid();
    NiftyReg::ConjugateGradient<float>::Perturbation(length);
    content.UpdateControlPointGrid();
    if (contentBw) contentBw->UpdateControlPointGrid();
    StoreCurrentDof();
}
