// A batch of F3D pairs evaluated together: every kernel of an objective or gradient evaluation
// runs once for all pairs (persistent sweeps over a (pair, tile) work list, item tables for the
// small kernels), on the batch's one stream.
#pragma once

#include "f3d_handle.hpp"

// Optimiser state of one pair of a batched registration (reg_base::Run / Optimiser::Optimise).
struct LsState {
    double best;                                        // best objective value
    float maxStep, smallest, currentSize, curLen, added;
    int iter, maxIter, lineIter, firstCall, justRestarted, done;
};

struct nrh_batch {
    std::vector<nrh_f3d *> pairs;
    std::vector<hipStream_t> ownStreams;  // each pair's own stream, restored when the batch ends
    bool ownsPairs = false;
    hipStream_t stream = nullptr;
    int P = 0;
    int device = 0;

    // fused sweeps (f3d_fused.hip): per-pair argument blocks and the (pair, tile) work list
    nrh::DevBuf<uint8_t> fusedArgs;   // FusedArgs[P]
    nrh::DevBuf<uint32_t> sweepItems;  // pair << 20 | tile, sorted by pair
    int nSweepItems = 0;
    int sweepBlocks = 0;               // persistent grid (a multiple of 8)
    size_t sweepLds = 0, fixLds = 0;
    nrh::DevBuf<int4> fixChunks;  // exact-pass chunks: (pair, first tile, tiles, -)
    int nFixChunks = 0;
    int rb = 0, fb = 0;
    nrh::DevBuf<uint8_t> projXArgs, projYArgs, nodeArgs;  // ProjXArgs[P], ProjYArgs[P], NodeArgs[P]
    nrh::DevBuf<int2> itProjX, itProjY, itNodes;
    int nProjX = 0, nProjY = 0, nNodes = 0;
    size_t projXLds = 0;

    // small kernels (f3d_ops.hip): per-pair descriptors and item tables (pair, block)
    nrh::DevBuf<uint8_t> pairOps;     // PairOps[P]
    nrh::DevBuf<int2> itBE, itG, itG3, itCg;
    int nBE = 0, nG = 0, nG3 = 0, nCg = 0;

    // results: per pair [NMI, H_r, H_f, H_rf, total, bending energy, value, max length]
    nrh::DevBuf<double> results;
    double *published = nullptr, *publishedDev = nullptr;  // coherent host copy of `results`
    nrh::DevBuf<float> weights;       // per pair [similarity, bending energy]

    // persistent sweep autotuning: workgroups per XCD, chosen by timing (0: not tuned yet)
    int tunedBlocksPerXcd = 0;
    double tuneMs[4] = { 0, 0, 0, 0 };

    // batched registration: per-pair optimiser state; flags [active, first, searching, store, -]
    nrh::DevBuf<LsState> lsState;
    nrh::DevBuf<int> lsFlags;
    nrh::DevBuf<float> lsScales, lsMaxStep;
    nrh::DevBuf<int> lsMaxIter;
    int *remaining = nullptr, *remainingDev = nullptr;  // pairs still iterating (coherent host memory)
};

namespace nrh {

constexpr int kBatchResults = 8;

// f3d_fused.hip
void batch_fused_setup(nrh_batch *b);
void batch_histogram(nrh_batch *b, const int *active);          // sweep + exact pass + finalize
void batch_similarity_gradient(nrh_batch *b, const int *active);  // K table, sweep, exact pass, projections
void batch_set_sweep_blocks(nrh_batch *b, int blocksPerXcd);
// f3d_ops.hip
void batch_ops_setup(nrh_batch *b);
void batch_finalize(nrh_batch *b, const int *active);
void batch_be_value(nrh_batch *b, const int *active);
void batch_be_gradient(nrh_batch *b, const int *active);       // with the max length of the result
void batch_normalise(nrh_batch *b, const int *active);
void batch_cg(nrh_batch *b, const int *active, const int *first);
void batch_update(nrh_batch *b, const int *active, const float *scales);
void batch_copy_dof(nrh_batch *b, const int *active, bool toBest);
void batch_values(nrh_batch *b, const int *active);             // results[] of the objective
// f3d_batch.hip
nrh_batch *batch_new(nrh_f3d *const *pairs, int n);
void batch_free(nrh_batch *b);
void batch_objective(nrh_batch *b, const int *active);
void batch_gradient(nrh_batch *b, const int *active, bool normalise);
void batch_touch(nrh_batch *b, bool pointsChanged);
void batch_autotune(nrh_batch *b);
void batch_ls_alloc(nrh_batch *b);
void batch_ls_init(nrh_batch *b, const std::vector<float>& maxStep, const std::vector<int>& maxIter);
void batch_ls_iteration(nrh_batch *b);  // launches only: capturable in a HIP graph

}  // namespace nrh
