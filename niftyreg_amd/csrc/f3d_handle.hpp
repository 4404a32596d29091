// State of one F3D registration pair on the device (the HIP counterpart of F3dContent).
#pragma once

#include <hip/hip_fp16.h>
#include "common.hpp"

struct nrh_f3d {
    nrh_geom ref{}, flo{}, cpp{};
    nrh_f3d_params params{};
    hipStream_t stream = nullptr;
    int device = 0;
    size_t N = 0, Nf = 0, G = 0;

    // images (float32 storage, or float16 storage when params.image_storage == 1)
    nrh::DevBuf<float> refImg, floImg;
    nrh::DevBuf<__half> refH, floH;
    nrh::DevBuf<int> mask;  // empty: every voxel active

    // control points (planar [3][G]) and optimiser arrays
    nrh::DevBuf<float> cppPos, best, nodeGrad, cgG, cgH;
    nrh::DevBuf<float> beGrad;  // bending-energy gradient sums computed beside the gradient sweep [3][G]

    // a second stream for work that overlaps the sweeps (the bending energy), forked from and
    // joined back to `stream` with the two events
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;

    // per-operation buffers, allocated on first use
    nrh::DevBuf<float> field, warped, warpedGrad, voxelGrad;
    nrh::DevBuf<float> convTmp, convDens, convDensTmp, convKernel;
    nrh::DevBuf<uint8_t> convNan;
    nrh::DevBuf<float> leDiag;      // per-node rotation-free displacement Jacobian diagonal [3][G]
    uint64_t leDiagVersion = 0;     // cppVersion leDiag was computed at (0: none)

    // joint histogram: integer counts, reg_nmi's Pro / Log arrays, entropies
    nrh::DevBuf<uint32_t> histCount;
    nrh::DevBuf<double> histPro, histLog;
    nrh::DevBuf<double> ent;       // H_r, H_f, H_rf, total, nmi
    nrh::DevBuf<float> histLogF;   // rb*fb joint log + rb + fb marginal logs, float

    // reduction scratch
    nrh::DevBuf<double> partials;  // per-block partial sums
    nrh::DevBuf<double> scalars;   // device-side scalar results
    nrh::DevBuf<uint32_t> maxBits;

    // fused node-gradient partial sums (one window per tile)

    // LNCC: two ping-pong sets of planar numerators + density, combined mask, reductions
    nrh::DevBuf<float> lnccA, lnccB, lnccKern;
    nrh::DevBuf<uint8_t> lnccComb;
    nrh::DevBuf<double> lnccPart, lnccScalars;  // scalars: adjusted weight, -, value, active count
    float lnccSigma = 0.f;
    int lnccRadius[3] = { 0, 0, 0 }, lnccKernOff[3] = { 0, 0, 0 };

    // velocity grid (reg_f3d2): intent_p1 SPLINE_VEL_GRID and intent_p2 (squaring steps; 0: choose)
    bool velocity = false;
    int velSteps = 0;
    nrh::DevBuf<float> velTmp;      // flow / ping-pong field [3][N]
    nrh::DevBuf<float> velFields;   // intermediate fields of the gradient exponentiation
    nrh::DevBuf<uint32_t> velKeys;  // min / max keys of the flow displacement
    hipEvent_t evPair = nullptr;    // cross-stream ordering of a forward / backward pair
    // the deformation field holds the exponentiation of the grid at velFieldVersion (0: it does
    // not) with velFieldSteps squarings; velFieldUpdated: the step number was chosen from the flow
    uint64_t velFieldVersion = 0;
    int velFieldSteps = 0;
    bool velFieldUpdated = false;
    nrh::DevBuf<float> velChain;      // the squaring's fields before the last, [k][3][N]
    uint64_t velChainVersion = 0;     // cppVersion of velChain (0: stale)
    // cppVersion at which the symmetric evaluation left the warped image and the joint histogram
    // of the current exponentiation (0: some other operation changed them since)
    uint64_t pairStateVersion = 0;
    // affine components of a velocity grid: its NIfTI extensions 0 (applied before the spline,
    // removed before the squaring and restored after it) and 1 (composed after it), and the
    // content's transformation matrix (removed from the opposite direction's intermediate fields)
    bool velExt0 = false, velExt1 = false, velTrans = false;
    float velExt0M[16], velExt1M[16], velTransM[16];

    double *pinned = nullptr;  // host-pinned scalar read-back
    double *published = nullptr, *publishedDev = nullptr;  // coherent host memory the device writes results to
    uint64_t publishSeq = 0;  // sequence number of the last publish (published[15] once it lands)

    uint64_t cppVersion = 1;   // bumped on every control-point change
    uint64_t histVersion = 0;  // cppVersion the histogram was computed at
    bool kTabCurrent = false;  // the last finalisation also wrote the fused sweep's NMI table
    bool halfDirty = true;     // float16 copies of the images need refreshing

    // optional HIP-event timing of the fused sweeps (on this handle's stream)
    bool profile = false;
    int profilePeriod = 1;        // time every profilePeriod-th sweep of each kind
    long profileSeen[2] = { 0, 0 };
    hipEvent_t ev[4] = { nullptr, nullptr, nullptr, nullptr };
    bool evPending[2] = { false, false };
    double msSweep[2] = { 0.0, 0.0 };  // accumulated: histogram sweep, gradient sweep
    long nSweep[2] = { 0, 0 };

    int totalBins() const { return params.ref_bins * params.flo_bins + params.ref_bins + params.flo_bins; }
};

namespace nrh {

void ensure_field(nrh_f3d *h);
void ensure_warped(nrh_f3d *h);
void ensure_warped_grad(nrh_f3d *h);
void ensure_voxel_grad(nrh_f3d *h);
const float *ref_ptr(nrh_f3d *h);  // float32 view (converted from float16 storage if needed)

// per-operation launches (f3d_ops.hip)
void op_spline_field(nrh_f3d *h);
void op_resample(nrh_f3d *h);
void op_image_gradient(nrh_f3d *h);
void op_joint_histogram(nrh_f3d *h, const float *warped);
// kTab (optional): also the fused gradient sweep's NMI table, when the histogram fits the LDS
// finalisation (h->kTabCurrent says whether it was written)
void op_nmi_finalize(nrh_f3d *h, float *kTab = nullptr, const double *pubS1 = nullptr, int pubN1 = -1);
void op_nmi_voxel_gradient(nrh_f3d *h, double weight);
void op_convolve_voxel_gradient(nrh_f3d *h);
void op_voxel_to_node(nrh_f3d *h, float weight);
// stream: where the launch goes (null: the handle's stream)
void op_bending_energy(nrh_f3d *h, double *out_device, hipStream_t stream = nullptr);
// the bending-energy gradient's gathered sums alone into h->beGrad (no weight, no round trip),
// for a projection that adds them (fused_gradient's BE overlap)
void op_bending_energy_gradient_raw(nrh_f3d *h, hipStream_t stream);
// trackMax: also leave the max length of the final gradient in maxBits[0] (float bits)
void op_bending_energy_gradient(nrh_f3d *h, float weight, bool trackMax = false);
void op_max_length(nrh_f3d *h);
// linear energy (value into out_device, / 3G; keepDiag marks the per-node diagonal as current) and
// its gradient (grad +=, optionally tracking the final max length in maxBits[0])
void op_linear_energy(nrh_f3d *h, double *out_device, bool keepDiag = true);
void op_linear_energy_gradient(nrh_f3d *h, float weight, bool trackMax = false);
void op_normalise(nrh_f3d *h, double max_length);
void op_rescale(nrh_f3d *h, float *data, size_t n, float lo, float hi);
void op_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
void op_fill_f32(float *p, size_t n, float v, hipStream_t s);
void op_cg(nrh_f3d *h, int first);
void op_cg_pair(nrh_f3d *fw, nrh_f3d *bw, int first);  // synchronous
void op_update_dof(nrh_f3d *h, float scale);
double read_scalar(nrh_f3d *h, const double *dev);
// Copies n0, n1, n2 doubles (at most 15 in all) from s0, s1, s2 (device) to h->published[0..]
// with one small kernel (no copy engine), then a new sequence number to published[15];
// wait_published returns once the host sees it (every earlier launch on the stream has finished).
void op_publish(nrh_f3d *h, const double *s0, int n0, const double *s1 = nullptr, int n1 = 0,
                const double *s2 = nullptr, int n2 = 0);
void wait_published(nrh_f3d *h);

// LNCC measure (lncc.hip); value = mean |local correlation| into out_device[0] (count in [1])
void op_lncc_value(nrh_f3d *h, double *out_device);
void op_lncc_voxel_gradient(nrh_f3d *h, double weight);

float node_weight_ratio(const nrh_f3d *h, float weight);
void op_normalise_device(nrh_f3d *h);
void op_max_length_bits(nrh_f3d *h);     // max length of nodeGrad into maxBits[0]
void op_normalise_from_bits(nrh_f3d *h);  // normalise by maxBits[0]; scalars[1] = the length

// velocity grids (f3d_velocity.hip)
void op_vel_def_field(nrh_f3d *h, bool update);
// syncOther: wait for bw's stream first (its grid may have been updated there)
void op_exponentiate_gradient(nrh_f3d *h, nrh_f3d *bw, bool syncOther = true);
void op_resample_gradient(nrh_f3d *h, float pad, float *out);
void op_compose_field(nrh_f3d *h, const nrh_geom& g, const float *def);
void op_update_velocity(nrh_f3d *h, float scale, bool ox, bool oy, bool oz);
void op_symmetrise_velocity(nrh_f3d *fw, nrh_f3d *bw);

// fused evaluation (f3d_fused.hip)
bool fused_supported(nrh_f3d *h);
void fused_release(const nrh_f3d *h);
// beValue: also the bending energy (into scalars[0]), in the exact pass's launch
// pubN1 >= 0: the finalisation also publishes the entropies and NMI followed by pubS1[0..pubN1)
// (op_publish's layout, the sequence number advanced)
void fused_histogram(nrh_f3d *h, float *fieldOut = nullptr, bool beValue = false, const double *pubS1 = nullptr,
                     int pubN1 = -1);
// beWeight > 0: the bending-energy gradient sums run on the side stream beside the x and y
// projection passes (op_bending_energy_gradient_raw); the node projection joins and adds them,
// round-trips the points and, with maxBits, tracks the final gradient's max length
void fused_gradient(nrh_f3d *h, float weight, float beWeight = 0.f, uint32_t *maxBits = nullptr);
// work on h->side after everything enqueued on h->stream so far / h->stream after h->side's work
void fork_side(nrh_f3d *h);
void join_side(nrh_f3d *h);
void collect_sweep_times(nrh_f3d *h);  // after a stream synchronisation
// The separable node projection of a voxel gradient (integer spacing, nodes on voxels): the
// convolution + VoxelCentricToNodeCentric of the reference as three passes, into nodeGrad.
bool fused_projection_supported(nrh_f3d *h);
void fused_project(nrh_f3d *h, const float *voxelGrad, float weight);

}  // namespace nrh
