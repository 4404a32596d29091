// HIP platform for NiftyReg's Platform / Content / Compute / Measure / Kernel / Optimiser
// interfaces, implemented over libnrhip's C ABI (include/niftyreg_hip.h).
//
// This is the reference-side half of the drop-in: a maintainer adds these files to reg-lib,
// a PlatformType::Hip value and its branch in Platform's constructor and CreateOptimiser
// (reg-lib/Platform.cpp:20-49,111-162), and links libnrhip.so.  Every class mirrors the CUDA
// platform's counterpart (reg-lib/cuda/Cuda*.h): the device keeps the images, control points
// and every intermediate; the Get*() accessors download into the host nifti images and the
// Update*() hooks upload, so host code that reads the content sees current data.
#pragma once

#include "AffineDeformationFieldKernel.h"
#include "AladinContent.h"
#include "AladinContentCreator.h"
#include "BlockMatchingKernel.h"
#include "ComputeFactory.h"
#include "ContentCreatorFactory.h"
#include "ConvolutionKernel.h"
#include "F3d2ContentCreator.h"
#include "F3dContent.h"
#include "F3dContentCreator.h"
#include "KernelFactory.h"
#include "LtsKernel.h"
#include "MeasureCreatorFactory.hpp"
#include "Optimiser.hpp"
#include "ResampleImageKernel.h"
#include "_reg_lncc.h"
#include "_reg_nmi.h"

#include "niftyreg_hip.h"

namespace NiftyRegHip {

// NR_FATAL_ERROR with libnrhip's message when rc != NRH_OK.
void Check(int rc, const char *call);
// Geometry of a nifti image as libnrhip reads it (sform if sform_code > 0, else qform).
nrh_geom GeomOf(const nifti_image *img);

}  // namespace NiftyRegHip

// ---------------------------------------------------------------- F3D content
class HipF3dContent: public F3dContent {
public:
    HipF3dContent(NiftiImage& referenceIn, NiftiImage& floatingIn, NiftiImage& controlPointGridIn,
                  NiftiImage *localWeightSimIn = nullptr, int *referenceMaskIn = nullptr,
                  mat44 *transformationMatrixIn = nullptr, size_t bytesIn = sizeof(float));
    virtual ~HipF3dContent();

    virtual NiftiImage& GetControlPointGrid() override;
    virtual NiftiImage& GetTransformationGradient() override;
    virtual NiftiImage& GetDeformationField() override;
    virtual NiftiImage& GetWarped() override;
    virtual NiftiImage& GetWarpedGradient() override;
    virtual NiftiImage& GetVoxelBasedMeasureGradient() override;
    virtual void UpdateControlPointGrid() override;
    virtual void UpdateTransformationGradient() override;
    virtual void ZeroTransformationGradient() override;
    virtual void ZeroVoxelBasedMeasureGradient() override;
    virtual void SetDeformationField(NiftiImage&& deformationFieldIn) override;
    virtual void UpdateDeformationField() override;
    virtual void UpdateVoxelBasedMeasureGradient() override;

    nrh_f3d* GetHandle() { return handle; }
    // The host grid's header (intent_p1 / intent_p2 of a velocity grid) without downloading the
    // control points.
    nifti_image* GridHeader() { return controlPointGrid; }
    // Marks the device control points as a velocity grid with the host header's step number
    // (reg_f3d2 keeps intent_p2 in the header; fails for grids with affine extensions).
    void SyncVelocityGrid();
    // Upload the host field / voxel gradient unless the device already holds the current one.
    void EnsureFieldOnDevice();
    void EnsureVoxelGradientOnDevice();
    // Called by the measures once their bin numbers / rescaling are known.
    void SetMeasure(int measure, int refBins, int floBins, float lnccSigma);
    // Device buffers that hold data (set by HipCompute and the measures); the accessors only
    // download those, so construction-time calls see the host images.
    enum Buffer { Field = 1, Warped = 2, WarpedGradient = 4, VoxelGradient = 8 };
    void MarkOnDevice(int buffers) { onDevice |= buffers; }

    // Fused route (environment variable NRH_PLUGIN_FUSED=1 at construction; F3D with a cubic
    // B-spline grid).  reg_f3d's stages of one objective value (field, warp, measure value) or one
    // similarity gradient (field, warp, image gradient, zeroed and measured voxel gradient,
    // convolution and node extraction) are recorded instead of run, and the measure value / the
    // convolution run them as one call (nrh_f3d_similarity_value / _gradient: the fused sweeps when
    // the geometry allows it).  A reader of any recorded buffer runs the recorded stages first with
    // the per-operation calls (Materialise) at the current control points: what the
    // per-operation route computes, unless the points changed since the stages were recorded (the
    // bending-energy gradient's round trip, an update), where the per-operation route would hold
    // the buffers of the old points.  reg_f3d reads them only after recomputing them.
    enum Stage { SField = 1, SWarp = 2, SImageGradient = 4, SZeroVoxel = 8, SVoxelGradient = 16, SConvolved = 32 };
    bool Fusing() const { return fusing; }
    int Pending() const { return pending; }
    void Defer(int stage) { pending |= stage; }
    void Materialise();
    void DropPending() { pending = 0; }  // a new chain of stages starts
    int Measure() const { return params.measure; }

private:
    nrh_f3d *handle = nullptr;
    nrh_f3d_params params{};
    int onDevice = 0;
    bool fusing = false;
    int pending = 0;
};

class HipF3dContentCreator: public F3dContentCreator {
public:
    virtual F3dContent* Create(NiftiImage& reference, NiftiImage& floating, NiftiImage& controlPointGrid,
                               NiftiImage *localWeightSim = nullptr, int *referenceMask = nullptr,
                               mat44 *transformationMatrix = nullptr, size_t bytes = sizeof(float)) override {
        return new HipF3dContent(reference, floating, controlPointGrid, localWeightSim, referenceMask,
                                 transformationMatrix, bytes);
    }
};

// ---------------------------------------------------------------- Aladin content
class HipAladinContent: public AladinContent {
public:
    HipAladinContent(NiftiImage& referenceIn, NiftiImage& floatingIn, int *referenceMaskIn = nullptr,
                     mat44 *transformationMatrixIn = nullptr, size_t bytesIn = sizeof(float),
                     const unsigned percentageOfBlocks = 0, const unsigned inlierLts = 0, int blockStepSize = 0);
    virtual ~HipAladinContent();

    virtual NiftiImage& GetWarped() override;
    virtual _reg_blockMatchingParam* GetBlockMatchingParams() override;

    nrh_aladin* GetHandle() { return handle; }

private:
    nrh_aladin *handle = nullptr;
};

// reg_f3d2's forward / backward pair (reg-lib/F3d2ContentCreator.h:6-21) as HIP contents.
class HipF3d2ContentCreator: public F3d2ContentCreator {
public:
    virtual std::pair<F3dContent*, F3dContent*> Create(NiftiImage& reference, NiftiImage& floating,
                                                       NiftiImage& controlPointGrid, NiftiImage& controlPointGridBw,
                                                       NiftiImage *localWeightSim = nullptr, int *referenceMask = nullptr,
                                                       int *floatingMask = nullptr, mat44 *transformationMatrix = nullptr,
                                                       mat44 *transformationMatrixBw = nullptr,
                                                       size_t bytes = sizeof(float)) override {
        auto con = new HipF3dContent(reference, floating, controlPointGrid, localWeightSim, referenceMask,
                                     transformationMatrix, bytes);
        auto conBw = new HipF3dContent(floating, reference, controlPointGridBw, nullptr, floatingMask,
                                       transformationMatrixBw, bytes);
        return { con, conBw };
    }
};

class HipAladinContentCreator: public AladinContentCreator {
public:
    virtual AladinContent* Create(NiftiImage& reference, NiftiImage& floating, int *referenceMask = nullptr,
                                  mat44 *transformationMatrix = nullptr, size_t bytes = sizeof(float),
                                  const unsigned percentageOfBlocks = 0, const unsigned inlierLts = 0,
                                  int blockStepSize = 0) override {
        return new HipAladinContent(reference, floating, referenceMask, transformationMatrix, bytes,
                                    percentageOfBlocks, inlierLts, blockStepSize);
    }
};

class HipContentCreatorFactory: public ContentCreatorFactory {
public:
    virtual ContentCreator* Produce(const ContentType conType) override;
};

// ---------------------------------------------------------------- Compute
class HipCompute: public Compute {
public:
    HipCompute(Content& con): Compute(con) {}
    virtual void ResampleImage(int interpolation, float paddingValue) override;
    virtual double ApproxBendingEnergy() override;
    virtual void ApproxBendingEnergyGradient(float weight) override;
    virtual double ApproxLinearEnergy() override;
    virtual void ApproxLinearEnergyGradient(float weight) override;
    virtual void GetDeformationField(bool composition, bool bspline) override;
    virtual void UpdateControlPointPosition(float *currentDof, const float *bestDof, const float *gradient,
                                            const float scale, const bool optimiseX, const bool optimiseY,
                                            const bool optimiseZ) override;
    virtual void GetImageGradient(int interpolation, float paddingValue, int activeTimePoint) override;
    virtual double GetMaximalLength(bool optimiseX, bool optimiseY, bool optimiseZ) override;
    virtual void NormaliseGradient(double maxGradLength, bool optimiseX, bool optimiseY, bool optimiseZ) override;
    virtual void ConvolveVoxelBasedMeasureGradient(float weight) override;
    virtual void GetAffineDeformationField(bool compose) override;
    virtual double GetJacobianPenaltyTerm(bool approx) override;
    virtual void JacobianPenaltyTermGradient(float weight, bool approx) override;
    virtual double CorrectFolding(bool approx) override;
    virtual double GetLandmarkDistance(size_t landmarkNumber, float *landmarkReference,
                                       float *landmarkFloating) override;
    virtual void LandmarkDistanceGradient(size_t landmarkNumber, float *landmarkReference, float *landmarkFloating,
                                          float weight) override;
    virtual void SmoothGradient(float sigma) override;
    virtual void GetApproximatedGradient(InterfaceOptimiser& opt) override;
    virtual void GetDefFieldFromVelocityGrid(const bool updateStepNumber) override;
    virtual void ExponentiateGradient(Content& conBw) override;
    virtual void UpdateVelocityField(float scale, bool optimiseX, bool optimiseY, bool optimiseZ) override;
    virtual void BchUpdate(float scale, int bchUpdateValue) override;
    virtual void SymmetriseVelocityFields(Content& conBw) override;
    virtual void DefFieldCompose(const NiftiImage& defField) override;
    virtual NiftiImage ResampleGradient(int interpolation, float padding) override;

protected:
    virtual void VoxelCentricToNodeCentric(float weight) override;

private:
    nrh_f3d* Handle();
};

class HipComputeFactory: public ComputeFactory {
public:
    virtual Compute* Produce(Content& con) override { return new HipCompute(con); }
};

// ---------------------------------------------------------------- Measures
class reg_nmi_hip: public reg_nmi {
public:
    // cBw: the backward content of a symmetric (reg_f3d2) registration
    void SetContent(HipF3dContent *c, HipF3dContent *cBw = nullptr) { content = c; contentBw = cBw; }
    virtual void InitialiseMeasure(nifti_image *refImg, nifti_image *floImg, int *refMask, nifti_image *warpedImg,
                                   nifti_image *warpedGrad, nifti_image *voxelBasedGrad,
                                   nifti_image *localWeightSim = nullptr, int *floMask = nullptr,
                                   nifti_image *warpedImgBw = nullptr, nifti_image *warpedGradBw = nullptr,
                                   nifti_image *voxelBasedGradBw = nullptr) override;
    virtual double GetSimilarityMeasureValueFw() override;
    virtual double GetSimilarityMeasureValueBw() override;
    virtual void GetVoxelBasedSimilarityMeasureGradientFw(int currentTimePoint) override;
    virtual void GetVoxelBasedSimilarityMeasureGradientBw(int currentTimePoint) override;

private:
    HipF3dContent *content = nullptr, *contentBw = nullptr;
};

class reg_lncc_hip: public reg_lncc {
public:
    // cBw: the backward content of a symmetric (reg_f3d2) registration
    void SetContent(HipF3dContent *c, HipF3dContent *cBw = nullptr) { content = c; contentBw = cBw; }
    virtual void InitialiseMeasure(nifti_image *refImg, nifti_image *floImg, int *refMask, nifti_image *warpedImg,
                                   nifti_image *warpedGrad, nifti_image *voxelBasedGrad,
                                   nifti_image *localWeightSim = nullptr, int *floMask = nullptr,
                                   nifti_image *warpedImgBw = nullptr, nifti_image *warpedGradBw = nullptr,
                                   nifti_image *voxelBasedGradBw = nullptr) override;
    virtual double GetSimilarityMeasureValueFw() override;
    virtual double GetSimilarityMeasureValueBw() override;
    virtual void GetVoxelBasedSimilarityMeasureGradientFw(int currentTimePoint) override;
    virtual void GetVoxelBasedSimilarityMeasureGradientBw(int currentTimePoint) override;

private:
    HipF3dContent *content = nullptr, *contentBw = nullptr;
};

class HipMeasureCreator: public MeasureCreator {
public:
    virtual reg_measure* Create(const MeasureType measureType) override;
    virtual void Initialise(reg_measure& measure, DefContent& con, DefContent *conBw = nullptr) override;
};

class HipMeasureCreatorFactory: public MeasureCreatorFactory {
public:
    virtual MeasureCreator* Produce() override { return new HipMeasureCreator(); }
};

// ---------------------------------------------------------------- Aladin kernels
class HipAffineDeformationFieldKernel: public AffineDeformationFieldKernel {
public:
    HipAffineDeformationFieldKernel(Content *con);
    virtual void Calculate(bool compose = false) override;

private:
    HipAladinContent *content;
};

class HipResampleImageKernel: public ResampleImageKernel {
public:
    HipResampleImageKernel(Content *con);
    virtual void Calculate(int interp, float paddingValue, bool *dtiTimePoint = nullptr,
                           mat33 *jacMat = nullptr) override;

private:
    HipAladinContent *content;
};

class HipBlockMatchingKernel: public BlockMatchingKernel {
public:
    HipBlockMatchingKernel(Content *con);
    virtual void Calculate() override;

private:
    HipAladinContent *content;
};

class HipLtsKernel: public LtsKernel {
public:
    HipLtsKernel(Content *con);
    virtual void Calculate(bool affine) override;

private:
    HipAladinContent *content;
};

class HipConvolutionKernel: public ConvolutionKernel {
public:
    virtual void Calculate(nifti_image *image, float *sigma, ConvKernelType kernelType, int *mask = nullptr,
                           bool *timePoints = nullptr, bool *axis = nullptr) override;
};

class HipKernelFactory: public KernelFactory {
public:
    virtual Kernel* Produce(std::string name, Content *con) const override;
};

// ---------------------------------------------------------------- Optimiser
// Conjugate gradient with the current / best / gradient arrays and the two CG arrays resident
// in the F3D handle (the HIP counterpart of CudaConjugateGradient, reg-lib/cuda/CudaOptimiser.hpp).
// With a backward content (reg_f3d2) both directions' arrays take part, as the reference's
// symmetric ConjugateGradient (reg-lib/Optimiser.cpp:231-310).
class HipConjugateGradient: public NiftyReg::ConjugateGradient<float> {
public:
    explicit HipConjugateGradient(HipF3dContent& con, HipF3dContent *conBw = nullptr): content(con), contentBw(conBw) {}
    virtual void StoreCurrentDof() override;
    virtual void RestoreBestDof() override;
    virtual void Perturbation(float length) override;
    virtual void RestartOptimisation() override { this->firstCall = true; }

protected:
    virtual void UpdateGradientValues() override;

private:
    HipF3dContent& content;
    HipF3dContent *contentBw;
};
