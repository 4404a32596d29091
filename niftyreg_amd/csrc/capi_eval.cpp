// Objective and gradient entry points of the F3D handle: the fused sweep when the geometry
// allows it, otherwise the per-operation chain of the reference's Compute calls.

#include "f3d_handle.hpp"

namespace nrh {

// reg_f3d::WarpFloatingImage + reg_nmi value (reg-lib/_reg_base.cpp:688-709, reg-lib/cpu/_reg_nmi.cpp:160-398)
static void histogram_at_current_points(nrh_f3d *h) {
    if (fused_supported(h)) {
        fused_histogram(h);
    } else {
        op_spline_field(h);
        op_resample(h);
        op_joint_histogram(h, h->warped.p);
        op_nmi_finalize(h);
        h->histVersion = h->cppVersion;
    }
}

}  // namespace nrh

using namespace nrh;

extern "C" {

// The objective's measure, bending and linear energies published to host memory.
static void publish_objective(nrh_f3d *h) {
    const bool lncc = h->params.measure == NRH_MEASURE_LNCC;
    op_publish(h, lncc ? h->scalars.p : h->ent.p, 5, h->scalars.p, 1, h->scalars.p + 9, 1);
}

// reg_f3d::GetObjectiveFunctionValue (reg-lib/_reg_f3d.cpp:443-463) enqueued: the measure, the
// bending and linear energies, and their publication to host memory.
static void enqueue_objective(nrh_f3d *h) {
    const nrh_f3d_params& p = h->params;
    const bool lncc = p.measure == NRH_MEASURE_LNCC;
    // the bending energy runs in the fused histogram's exact-pass launch
    const bool beFused = p.be_weight > 0 && p.similarity_weight > 0 && !lncc && fused_supported(h);
    if (p.similarity_weight > 0) {
        if (lncc) {
            // reg_f3d::WarpFloatingImage then reg_lncc::GetSimilarityMeasureValueFw
            op_spline_field(h);
            op_resample(h);
            op_lncc_value(h, h->scalars.p + 4);
        } else if (beFused) {
            // with no linear energy to wait for, the finalisation publishes the results itself
            if (p.le_weight > 0) fused_histogram(h, nullptr, true);
            else fused_histogram(h, nullptr, true, h->scalars.p, 1);
        } else {
            histogram_at_current_points(h);
        }
    }
    const bool published = beFused && p.le_weight <= 0;
    if (p.be_weight > 0 && !beFused) op_bending_energy(h, h->scalars.p + 0);
    if (p.le_weight > 0) op_linear_energy(h, h->scalars.p + 9);
    if (!published) publish_objective(h);
}

// The published objective: value = wMeasure - wBE - wLE, terms (may be null) the three terms.
static void finish_objective(nrh_f3d *h, double *value, double terms[3]) {
    const nrh_f3d_params& p = h->params;
    wait_published(h);
    collect_sweep_times(h);
    const double *r = h->published;
    const double wMeasure = p.similarity_weight > 0 ? (double)p.similarity_weight * r[4] : 0.0;
    const double wBE = p.be_weight > 0 ? (double)p.be_weight * r[5] : 0.0;
    const double wLE = p.le_weight > 0 ? (double)p.le_weight * r[6] : 0.0;
    *value = wMeasure - wBE - wLE;
    if (terms) {
        terms[0] = wMeasure;
        terms[1] = wBE;
        terms[2] = wLE;
    }
}

int nrh_f3d_objective(nrh_f3d *h, double *value, double terms[3]) {
    return guard([&] {
        NRH_REQUIRE(h && value, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(!h->velocity, NRH_ERR_UNSUPPORTED, "velocity grids are evaluated with nrh_f3d2_objective");
        enqueue_objective(h);
        finish_objective(h, value, terms);
    });
}

int nrh_f3d_similarity_value(nrh_f3d *h, double *value) {
    return guard([&] {
        NRH_REQUIRE(h && value, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(!h->velocity, NRH_ERR_UNSUPPORTED, "velocity grids are evaluated with nrh_f3d2_objective");
        const bool lncc = h->params.measure == NRH_MEASURE_LNCC;
        if (lncc) {
            // reg_f3d::WarpFloatingImage then reg_lncc::GetSimilarityMeasureValueFw
            op_spline_field(h);
            op_resample(h);
            op_lncc_value(h, h->scalars.p + 4);
        } else {
            histogram_at_current_points(h);
        }
        op_publish(h, lncc ? h->scalars.p : h->ent.p, 5, nullptr, 0, nullptr, 0);
        wait_published(h);
        collect_sweep_times(h);
        *value = h->published[4];
    });
}

int nrh_f3d_similarity_gradient(nrh_f3d *h, float weight) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        NRH_REQUIRE(!h->velocity, NRH_ERR_UNSUPPORTED, "velocity grids are evaluated with nrh_f3d2_gradient");
        if (h->params.measure == NRH_MEASURE_LNCC) {
            op_spline_field(h);
            op_resample(h);
            op_image_gradient(h);
            ensure_voxel_grad(h);
            NRH_HIP(hipMemsetAsync(h->voxelGrad.p, 0, h->voxelGrad.bytes(), h->stream));
            op_lncc_voxel_gradient(h, 1.0);
            if (fused_projection_supported(h)) {
                fused_project(h, h->voxelGrad.p, weight);
            } else {
                op_convolve_voxel_gradient(h);
                op_voxel_to_node(h, weight);
            }
        } else if (fused_supported(h)) {
            if (h->histVersion != h->cppVersion) fused_histogram(h);
            fused_gradient(h, weight, 0.f, nullptr);
        } else {
            histogram_at_current_points(h);
            op_image_gradient(h);
            ensure_voxel_grad(h);
            NRH_HIP(hipMemsetAsync(h->voxelGrad.p, 0, h->voxelGrad.bytes(), h->stream));
            op_nmi_voxel_gradient(h, 1.0);
            op_convolve_voxel_gradient(h);
            op_voxel_to_node(h, weight);
        }
    });
}

// reg_f3d::GetObjectiveFunctionGradient (reg-lib/_reg_f3d.cpp:553-577); with max_length, normalised
// and the length returned (a synchronisation), else only enqueued.
static void gradient_body(nrh_f3d *h, double *max_length) {
        const nrh_f3d_params& p = h->params;
        const bool be = p.be_weight > 0, le = p.le_weight > 0;
        bool beDone = false;
        if (p.similarity_weight > 0 && p.measure == NRH_MEASURE_LNCC) {
            // reg_f3d::GetSimilarityMeasureGradient with reg_lncc (reg-lib/_reg_f3d.cpp:353-362,
            // reg-lib/_reg_base.cpp:497-549)
            op_spline_field(h);
            op_resample(h);
            op_image_gradient(h);
            ensure_voxel_grad(h);
            NRH_HIP(hipMemsetAsync(h->voxelGrad.p, 0, h->voxelGrad.bytes(), h->stream));
            op_lncc_voxel_gradient(h, 1.0);
            // convolution + node extraction: the separable projection when the nodes lie on
            // voxels at an integer spacing, else the reference-order convolution
            if (fused_projection_supported(h)) {
                fused_project(h, h->voxelGrad.p, p.similarity_weight);
            } else {
                op_convolve_voxel_gradient(h);
                op_voxel_to_node(h, p.similarity_weight);
            }
        } else if (p.similarity_weight > 0) {
            if (fused_supported(h)) {
                // the bending-energy gradient sums run on the side stream beside the projection
                // passes; the node projection adds them
                const bool trackMax = be && max_length != nullptr && !le;
                if (trackMax) op_fill_u32(h->maxBits.p, 1, 0u, h->stream);
                if (h->histVersion != h->cppVersion) fused_histogram(h);
                fused_gradient(h, p.similarity_weight, be ? p.be_weight : 0.f, trackMax ? h->maxBits.p : nullptr);
                beDone = be;
            } else {
                // reg_f3d::GetSimilarityMeasureGradient chain (reg-lib/_reg_f3d.cpp:353-362)
                op_spline_field(h);
                op_resample(h);
                op_joint_histogram(h, h->warped.p);
                op_nmi_finalize(h);
                h->histVersion = h->cppVersion;
                op_image_gradient(h);
                ensure_voxel_grad(h);
                NRH_HIP(hipMemsetAsync(h->voxelGrad.p, 0, h->voxelGrad.bytes(), h->stream));
                op_nmi_voxel_gradient(h, 1.0);
                op_convolve_voxel_gradient(h);
                op_voxel_to_node(h, p.similarity_weight);
            }
        } else {
            op_fill_f32(h->nodeGrad.p, 3 * h->G, 0.f, h->stream);
        }
        // reg_f3d::GetObjectiveFunctionGradient: bending energy, then linear energy
        if (be && !beDone) op_bending_energy_gradient(h, p.be_weight, max_length != nullptr && !le);
        if (le) op_linear_energy_gradient(h, p.le_weight, max_length != nullptr);
        if (max_length) {
            if (!be && !le) op_max_length_bits(h);
            op_normalise_from_bits(h);
            *max_length = read_scalar(h, h->scalars.p + 1);
            collect_sweep_times(h);
        }
        // without a length to return the call only enqueues: the gradient stays on the device,
        // and the downloads (or the next call that returns a value) wait for it
}

int nrh_f3d_gradient(nrh_f3d *h, double *max_length) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        NRH_REQUIRE(!h->velocity, NRH_ERR_UNSUPPORTED, "velocity grids are evaluated with nrh_f3d2_gradient");
        gradient_body(h, max_length);
    });
}

int nrh_f3d_set_profiling(nrh_f3d *h, int enable) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        if (enable && !h->ev[0])
            for (int i = 0; i < 4; ++i) NRH_HIP(hipEventCreate(&h->ev[i]));
        NRH_REQUIRE(enable >= 0, NRH_ERR_ARG, "the timing period must be >= 0");
        h->profile = enable != 0;
        h->profilePeriod = enable > 0 ? enable : 1;
        h->profileSeen[0] = h->profileSeen[1] = 0;
        h->msSweep[0] = h->msSweep[1] = 0.0;
        h->nSweep[0] = h->nSweep[1] = 0;
        h->evPending[0] = h->evPending[1] = false;
    });
}

int nrh_f3d_sweep_times(nrh_f3d *h, double ms[2], long counts[2]) {
    return guard([&] {
        NRH_REQUIRE(h && ms && counts, NRH_ERR_ARG, "null argument");
        NRH_HIP(hipStreamSynchronize(h->stream));
        collect_sweep_times(h);
        for (int k = 0; k < 2; ++k) {
            ms[k] = h->msSweep[k];
            counts[k] = h->nSweep[k];
        }
    });
}

int nrh_f3d_fused_field(nrh_f3d *h) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        NRH_REQUIRE(fused_supported(h), NRH_ERR_UNSUPPORTED, "fused path unavailable for this geometry");
        ensure_field(h);
        h->velFieldVersion = 0;
        fused_histogram(h, h->field.p);
        NRH_HIP(hipStreamSynchronize(h->stream));
        collect_sweep_times(h);
    });
}

int nrh_f3d_fused_available(nrh_f3d *h, int *available) {
    return guard([&] {
        NRH_REQUIRE(h && available, NRH_ERR_ARG, "null argument");
        *available = fused_supported(h) ? 1 : 0;
    });
}

}  // extern "C"
