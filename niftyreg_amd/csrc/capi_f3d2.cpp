// Symmetric diffeomorphic F3D (reg_f3d2) objective and gradient over two handles: the forward
// one (reference -> floating, forward velocity grid) and the backward one (floating ->
// reference, backward velocity grid, bin counts swapped).  Each handle's chain runs on its own
// stream, so the two directions overlap on the device; they meet where the reference couples
// them: the backward grid takes the forward step count, each gradient is exponentiated along
// the other direction's grid (both before either bending-energy gradient rounds its grid
// through displacements), and the two gradients share one normalisation length.

#include "f3d_handle.hpp"
#include <algorithm>

namespace nrh {

namespace {

void check_pair(nrh_f3d *fw, nrh_f3d *bw) {
    NRH_REQUIRE(fw && bw && fw != bw, NRH_ERR_ARG, "two distinct handles are required");
    NRH_REQUIRE(fw->velocity && bw->velocity, NRH_ERR_STATE, "both control point grids must be velocity grids");
    const nrh_f3d_params &p = fw->params, &q = bw->params;
    NRH_REQUIRE(p.measure == q.measure, NRH_ERR_ARG, "forward and backward handles must use the same measure");
    NRH_REQUIRE(p.measure == NRH_MEASURE_NMI || p.lncc_sigma == q.lncc_sigma, NRH_ERR_ARG,
                "forward and backward LNCC kernels differ");
    NRH_REQUIRE(p.similarity_weight == q.similarity_weight && p.be_weight == q.be_weight && p.le_weight == q.le_weight,
                NRH_ERR_ARG, "forward and backward handles must share the objective weights");
    NRH_REQUIRE(p.ref_bins == q.flo_bins && p.flo_bins == q.ref_bins, NRH_ERR_ARG,
                "the backward handle's bin counts must be the forward ones swapped");
}

hipEvent_t pair_event(nrh_f3d *h) {
    if (!h->evPair) NRH_HIP(hipEventCreateWithFlags(&h->evPair, hipEventDisableTiming));
    return h->evPair;
}

// Each stream waits for the other's work issued so far.
void join(nrh_f3d *a, nrh_f3d *b) {
    NRH_HIP(hipEventRecord(pair_event(a), a->stream));
    NRH_HIP(hipEventRecord(pair_event(b), b->stream));
    NRH_HIP(hipStreamWaitEvent(a->stream, b->evPair, 0));
    NRH_HIP(hipStreamWaitEvent(b->stream, a->evPair, 0));
}

// reg_f3d2::GetDeformationField + WarpFloatingImage (reg-lib/_reg_f3d2.cpp:139-172) and the
// symmetric NMI's joint histograms (reg-lib/cpu/_reg_measure.h:63-80): forward exponentiation
// (step number updated: the optimiser exists), the backward one with the forward step count,
// both resamplings and histograms.  A gradient at the control points of the last objective
// finds all of it in place: the reference recomputes the same values.
void warp_and_histograms(nrh_f3d *fw, nrh_f3d *bw) {
    op_vel_def_field(fw, true);
    bw->velSteps = fw->velSteps;
    op_vel_def_field(bw, false);
    for (nrh_f3d *h : { fw, bw }) {
        if (h->pairStateVersion == h->cppVersion) continue;
        op_resample(h);
        if (h->params.measure == NRH_MEASURE_NMI) {
            op_joint_histogram(h, h->warped.p);
            op_nmi_finalize(h);
            h->histVersion = h->cppVersion;
        }
        h->pairStateVersion = h->cppVersion;
    }
}

}  // namespace

}  // namespace nrh

using namespace nrh;

extern "C" {

// reg_f3d2::GetObjectiveFunctionValue (reg-lib/_reg_f3d2.cpp:540-560) with NMI, bending and
// linear energy: wMeasure = w_sim (NMI_fw + NMI_bw), wBE = w_be BE_fw + w_be BE_bw, likewise wLE.
int nrh_f3d2_objective(nrh_f3d *fw, nrh_f3d *bw, double *value, double terms[3]) {
    return guard([&] {
        check_pair(fw, bw);
        NRH_REQUIRE(value, NRH_ERR_ARG, "null argument");
        const nrh_f3d_params& p = fw->params;
        const bool lncc = p.measure == NRH_MEASURE_LNCC;
        if (p.similarity_weight > 0) {
            warp_and_histograms(fw, bw);
            // reg_lncc::GetSimilarityMeasureValueFw / Bw (reg-lib/cpu/_reg_lncc.cpp:357-390)
            if (lncc)
                for (nrh_f3d *h : { fw, bw }) op_lncc_value(h, h->scalars.p + 4);
        }
        for (nrh_f3d *h : { fw, bw }) {
            if (p.be_weight > 0) op_bending_energy(h, h->scalars.p + 0);
            if (p.le_weight > 0) op_linear_energy(h, h->scalars.p + 9);
            op_publish(h, lncc ? h->scalars.p : h->ent.p, 5, h->scalars.p, 1, h->scalars.p + 9, 1);
        }
        wait_published(fw);
        wait_published(bw);
        const double *f = fw->published, *b = bw->published;
        const double wMeasure = p.similarity_weight > 0 ? (double)p.similarity_weight * (f[4] + b[4]) : 0.0;
        const double wBE = p.be_weight > 0 ? (double)p.be_weight * f[5] + (double)p.be_weight * b[5] : 0.0;
        const double wLE = p.le_weight > 0 ? (double)p.le_weight * f[6] + (double)p.le_weight * b[6] : 0.0;
        *value = wMeasure - wBE - wLE;
        if (terms) {
            terms[0] = wMeasure;
            terms[1] = wBE;
            terms[2] = wLE;
        }
    });
}

// reg_f3d2::GetObjectiveFunctionGradient (reg-lib/_reg_f3d2.cpp:417-440) and, when max_length
// is non-NULL, reg_f3d2::NormaliseGradient (:373-395): both node gradients divided by the larger
// of the two maximal lengths.
int nrh_f3d2_gradient(nrh_f3d *fw, nrh_f3d *bw, double *max_length) {
    return guard([&] {
        check_pair(fw, bw);
        const nrh_f3d_params& p = fw->params;
        if (p.similarity_weight > 0) {
            // reg_f3d2::GetVoxelBasedGradient (reg-lib/_reg_f3d2.cpp:246-301)
            warp_and_histograms(fw, bw);
            for (nrh_f3d *h : { fw, bw }) {
                ensure_voxel_grad(h);
                op_fill_f32(h->voxelGrad.p, 3 * h->N, 0.f, h->stream);
                op_image_gradient(h);
            }
            for (nrh_f3d *h : { fw, bw }) {
                if (p.measure == NRH_MEASURE_LNCC) op_lncc_voxel_gradient(h, 1.0);
                else op_nmi_voxel_gradient(h, 1.0);
            }
            // ExponentiateGradient (reg-lib/_reg_f3d2.cpp:597-610): each direction reads the
            // other's grid (unchanged until the join before the penalty gradients) and, on a
            // shared geometry, the other's squaring chain, written on the other stream
            join(fw, bw);
            op_exponentiate_gradient(fw, bw, false);
            op_exponentiate_gradient(bw, fw, false);
            // reg_f3d2::GetSimilarityMeasureGradient (:333-342): convolution and node extraction,
            // as the separable projection when the grid's nodes lie on voxels at an integer spacing
            for (nrh_f3d *h : { fw, bw }) {
                if (fused_projection_supported(h)) {
                    fused_project(h, h->voxelGrad.p, p.similarity_weight);
                } else {
                    op_convolve_voxel_gradient(h);
                    op_voxel_to_node(h, p.similarity_weight);
                }
            }
        } else {
            op_fill_f32(fw->nodeGrad.p, 3 * fw->G, 0.f, fw->stream);
            op_fill_f32(bw->nodeGrad.p, 3 * bw->G, 0.f, bw->stream);
        }
        // the bending-energy gradient rounds each grid through displacements in place: only after
        // both exponentiations have read the grids
        join(fw, bw);
        for (nrh_f3d *h : { fw, bw }) {
            if (p.be_weight > 0) op_bending_energy_gradient(h, p.be_weight);
            if (p.le_weight > 0) op_linear_energy_gradient(h, p.le_weight);
        }
        if (max_length) {
            op_max_length(fw);
            op_max_length(bw);
            const float lf = (float)read_scalar(fw, fw->scalars.p + 1);
            const float lb = (float)read_scalar(bw, bw->scalars.p + 1);
            const float m = std::max(lb, lf);
            op_normalise(fw, m);
            op_normalise(bw, m);
            *max_length = m;
        }
        NRH_HIP(hipStreamSynchronize(fw->stream));
        NRH_HIP(hipStreamSynchronize(bw->stream));
    });
}

// ConjugateGradient<T>::UpdateGradientValues with a backward direction (reg-lib/Optimiser.cpp:
// 231-310): one gamma from both directions' dot products.
int nrh_f3d2_conjugate_gradient(nrh_f3d *fw, nrh_f3d *bw, int first) {
    return guard([&] {
        NRH_REQUIRE(fw && bw && fw != bw, NRH_ERR_ARG, "two distinct handles are required");
        op_cg_pair(fw, bw, first);
    });
}

}  // extern "C"
