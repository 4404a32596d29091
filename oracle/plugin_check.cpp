// Drive the reference's own F3D evaluation calls through either platform (test infrastructure).
//
// Compiled by oracle/Makefile (`make plugin`) with the reference's sources' headers, linked to
// the scalar reference build (oracle/_ref/libnrref_nosse.so) and to libnrhip.so.  The same
// sequence of Content / Compute / Measure calls that reg_f3d makes for one objective value and
// one objective gradient (reg-lib/_reg_f3d.cpp:443-463,553-577, reg-lib/_reg_base.cpp:497-549)
// runs on the reference's CPU classes (hip = 0) or on the HIP platform classes of
// niftyreg_amd/plugin (hip = 1), so a test can compare the two results.

#include "HipPlatform.hpp"
#include "ref_nifti.hpp"

#include <limits>
#include <memory>

using namespace ref_nifti;

namespace {
thread_local std::string plugin_error;
}

extern "C" {

const char* plugin_last_error() { return plugin_error.c_str(); }

// measure = similarity value, be = bending energy + 1000 x linear energy (leW > 0); grad =
// transformation gradient (3 G) after the similarity, bending-energy and linear-energy (weight
// leW, skipped when 0) terms; voxelGrad (3 N, may be NULL) = convolved voxel gradient.
int plugin_f3d_eval(int hip, const nr_geom *ref, const float *refData, const nr_geom *flo, const float *floData,
                    const nr_geom *cpp, const float *cppData, int bins, int lncc, float simW, float beW,
                    float leW, double *measure, double *be, float *grad, float *voxelGrad) {
    try {
        NiftiImage r = make_image(ref, refData), f = make_image(flo, floData), g = make_grid(cpp, cppData);
        std::unique_ptr<F3dContentCreator> cc(hip ? new HipF3dContentCreator() : new F3dContentCreator());
        std::unique_ptr<F3dContent> con(cc->Create(r, f, g));
        std::unique_ptr<ComputeFactory> cf(hip ? new HipComputeFactory() : new ComputeFactory());
        std::unique_ptr<Compute> comp(cf->Produce(*con));
        std::unique_ptr<MeasureCreator> mc(hip ? new HipMeasureCreator() : new MeasureCreator());
        std::unique_ptr<reg_measure> m(mc->Create(lncc ? MeasureType::Lncc : MeasureType::Nmi));
        if (lncc) {
            dynamic_cast<reg_lncc&>(*m).SetKernelStandardDeviation(0, -5.f);
        } else {
            dynamic_cast<reg_nmi&>(*m).SetReferenceBinNumber(bins + 4, 0);
            dynamic_cast<reg_nmi&>(*m).SetFloatingBinNumber(bins + 4, 0);
        }
        m->SetTimePointWeight(0, 1.0);
        mc->Initialise(*m, *con);
        const float pad = std::numeric_limits<float>::quiet_NaN();
        // objective value
        comp->GetDeformationField(false, true);
        comp->ResampleImage(1, pad);
        *measure = m->GetSimilarityMeasureValue();
        *be = comp->ApproxBendingEnergy() + (leW > 0 ? 1000.0 * comp->ApproxLinearEnergy() : 0.0);
        // objective gradient
        comp->GetDeformationField(false, true);
        comp->ResampleImage(1, pad);
        comp->GetImageGradient(1, pad, 0);
        con->ZeroVoxelBasedMeasureGradient();
        m->GetVoxelBasedSimilarityMeasureGradient(0);
        comp->ConvolveVoxelBasedMeasureGradient(simW);
        comp->ApproxBendingEnergyGradient(beW);
        if (leW > 0) comp->ApproxLinearEnergyGradient(leW);
        NiftiImage& tg = con->GetTransformationGradient();
        std::memcpy(grad, tg->data, tg.nVoxels() * sizeof(float));
        if (voxelGrad) {
            NiftiImage& vg = con->GetVoxelBasedMeasureGradient();
            std::memcpy(voxelGrad, vg->data, vg.nVoxels() * sizeof(float));
        }
        return 0;
    } catch (const std::exception& e) {
        plugin_error = e.what();
        return 1;
    }
}

}  // extern "C"
