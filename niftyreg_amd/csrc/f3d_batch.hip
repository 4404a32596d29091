// Batches of F3D pairs: the C ABI of a batch (many reg_f3d pairs of different sizes evaluated by
// the same launches) and the device-side conjugate-gradient / line-search iteration of
// reg_base<T>::Run (reg-lib/_reg_base.cpp:727-819, reg-lib/Optimiser.cpp:129-173) for every pair
// at once, which a registration of the batch captures once per level as a HIP graph and replays.
//
// Each pair keeps its own optimiser state on the device; an iteration runs the gradient of every
// pair still iterating, then the twelve line-search steps of Optimiser::Optimise in lockstep: a
// step updates, evaluates and judges only the pairs whose own line search is still running, so
// every pair takes exactly the decisions it takes alone (the objective values are bit-identical
// to a pair's own evaluation).

#include "f3d_batch.hpp"
#include <algorithm>
#include <memory>

namespace nrh {

namespace {

__global__ void k_ls_init(LsState *__restrict__ st, int n, const double *__restrict__ res, const float *__restrict__ maxStep,
                          const int *__restrict__ maxIter) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    LsState s{};
    s.best = res[q * kBatchResults + 6];
    s.maxStep = maxStep[q];
    s.smallest = maxStep[q] / 100.f;
    s.currentSize = maxStep[q];
    s.maxIter = maxIter[q];
    s.firstCall = 1;
    s.done = maxIter[q] <= 0;
    st[q] = s;
}

// Top of reg_base::Run's loop for every pair: stop at the iteration budget, else count the
// gradient evaluation, cap the step and start the line search.
__global__ void k_ls_begin(LsState *__restrict__ st, int n, int *__restrict__ active, int *__restrict__ first,
                           int *__restrict__ searching, float *__restrict__ scales) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    LsState s = st[q];
    if (!s.done && s.iter >= s.maxIter) s.done = 1;
    active[q] = !s.done;
    searching[q] = 0;
    if (s.done) {
        st[q] = s;
        return;
    }
    ++s.iter;
    s.currentSize = fminf(s.currentSize, s.maxStep);
    first[q] = s.firstCall;
    s.firstCall = 0;
    s.lineIter = 0;
    s.added = 0.f;
    s.curLen = s.currentSize;
    searching[q] = s.curLen > s.smallest && s.lineIter < 12 && s.iter < s.maxIter;
    scales[q] = -s.curLen;
    st[q] = s;
}

// One line-search step's decision (Optimiser::Optimise): keep an improving step and grow the next
// one (x1.1, capped), else halve it; store[q] asks for the current points to be kept as the best.
__global__ void k_ls_decide(LsState *__restrict__ st, int n, int *__restrict__ searching, int *__restrict__ store,
                            float *__restrict__ scales, const double *__restrict__ res) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    store[q] = 0;
    if (!searching[q]) return;
    LsState s = st[q];
    const double value = res[q * kBatchResults + 6];
    if (value > s.best) {
        s.best = value;
        s.added += s.curLen;
        s.curLen *= 1.1f;
        s.curLen = fminf(s.curLen, s.maxStep);
        store[q] = 1;
    } else {
        s.curLen *= 0.5f;
    }
    ++s.iter;
    ++s.lineIter;
    searching[q] = s.curLen > s.smallest && s.lineIter < 12 && s.iter < s.maxIter;
    scales[q] = -s.curLen;
    st[q] = s;
}

// End of the iteration: the line search's total step becomes the next start size; a stalled
// search restarts once from steepest descent at the full step (reg-lib/_reg_base.cpp:781-793),
// otherwise the pair is done.  remaining: the number of pairs still iterating.
__global__ void k_ls_end(LsState *__restrict__ st, int n, const int *__restrict__ active, int *__restrict__ remaining) {
    __shared__ int count;
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        LsState s = st[q];
        if (active[q]) {
            s.currentSize = s.added;
            if (s.currentSize == 0.f) {
                if (!s.justRestarted && s.iter < s.maxIter) {
                    s.firstCall = 1;
                    s.currentSize = s.maxStep;
                    s.justRestarted = 1;
                }
            } else {
                s.justRestarted = 0;
            }
            if (s.currentSize == 0.f) s.done = 1;
            st[q] = s;
        }
        if (!s.done && s.iter < s.maxIter) atomicAdd(&count, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *remaining = count;
}

}  // namespace

void batch_objective(nrh_batch *b, const int *active) {
    batch_histogram(b, active);
    batch_be_value(b, active);
    batch_values(b, active);
}

// nrh_f3d_gradient(h, &max_length) for every active pair: the histogram at the current points,
// the similarity gradient, the bending-energy gradient (with the displacement round trip of the
// points) and the normalisation by the max length.
void batch_gradient(nrh_batch *b, const int *active, bool normalise) {
    batch_histogram(b, active);
    batch_similarity_gradient(b, active);
    batch_be_gradient(b, active);
    if (normalise) batch_normalise(b, active);
    batch_values(b, active);
}

void batch_ls_alloc(nrh_batch *b) {
    b->lsState.alloc(b->P);
    b->lsFlags.alloc(5 * (size_t)b->P);
    b->lsScales.alloc(b->P);
    b->lsMaxStep.alloc(b->P);
    b->lsMaxIter.alloc(b->P);
    if (!b->remaining) {
        NRH_HIP(hipHostMalloc((void **)&b->remaining, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
        NRH_HIP(hipHostGetDevicePointer((void **)&b->remainingDev, b->remaining, 0));
    }
}

void batch_ls_init(nrh_batch *b, const std::vector<float>& maxStep, const std::vector<int>& maxIter) {
    NRH_HIP(hipMemcpyAsync(b->lsMaxStep.p, maxStep.data(), b->P * sizeof(float), hipMemcpyHostToDevice, b->stream));
    NRH_HIP(hipMemcpyAsync(b->lsMaxIter.p, maxIter.data(), b->P * sizeof(int), hipMemcpyHostToDevice, b->stream));
    // Optimiser::Initialise: the current points are the best ones, and their objective the best value
    batch_copy_dof(b, nullptr, true);
    batch_objective(b, nullptr);
    const unsigned g = (unsigned)((b->P + 63) / 64);
    hipLaunchKernelGGL(k_ls_init, dim3(g), dim3(64), 0, b->stream, b->lsState.p, b->P, b->results.p, b->lsMaxStep.p,
                       b->lsMaxIter.p);
    NRH_HIP(hipGetLastError());
}

// One iteration of reg_base::Run for every pair still iterating: launches only (capturable).
void batch_ls_iteration(nrh_batch *b) {
    const unsigned g = (unsigned)((b->P + 63) / 64);
    int *active = b->lsFlags.p, *first = active + b->P, *searching = first + b->P, *store = searching + b->P;
    hipLaunchKernelGGL(k_ls_begin, dim3(g), dim3(64), 0, b->stream, b->lsState.p, b->P, active, first, searching,
                       b->lsScales.p);
    batch_gradient(b, active, true);
    batch_cg(b, active, first);
    for (int step = 0; step < 12; ++step) {
        batch_update(b, searching, b->lsScales.p);
        batch_objective(b, searching);
        hipLaunchKernelGGL(k_ls_decide, dim3(g), dim3(64), 0, b->stream, b->lsState.p, b->P, searching, store,
                           b->lsScales.p, b->results.p);
        batch_copy_dof(b, store, true);
    }
    batch_copy_dof(b, active, false);
    hipLaunchKernelGGL(k_ls_end, dim3(1), dim3(1024), 0, b->stream, b->lsState.p, b->P, active, b->remainingDev);
    NRH_HIP(hipGetLastError());
}

// The per-pair handles follow the batch's work: pointsChanged (the bending-energy round trip or a
// line search moved them) invalidates what the handles keep of their points; otherwise the
// histogram is that of the current points.
void batch_touch(nrh_batch *b, bool pointsChanged) {
    for (nrh_f3d *h : b->pairs) {
        if (pointsChanged) {
            h->cppVersion++;
            h->leDiagVersion = 0;
        } else {
            h->histVersion = h->cppVersion;
        }
        h->kTabCurrent = false;  // the batch's finalisation leaves the pair's own table as it was
        h->pairStateVersion = 0;
        h->velFieldVersion = 0;
    }
}

// Times the persistent histogram sweep of the batch for a few grid sizes (workgroups per XCD)
// and keeps the fastest.
void batch_autotune(nrh_batch *b) {
    const int cand[4] = { 32, 64, 128, 256 };
    hipEvent_t e0, e1;
    NRH_HIP(hipEventCreate(&e0));
    NRH_HIP(hipEventCreate(&e1));
    int bestK = 128;
    double bestMs = 1e30;
    for (int i = 0; i < 4; ++i) {
        batch_set_sweep_blocks(b, cand[i]);
        batch_histogram(b, nullptr);  // warm-up
        NRH_HIP(hipEventRecord(e0, b->stream));
        for (int r = 0; r < 3; ++r) batch_histogram(b, nullptr);
        NRH_HIP(hipEventRecord(e1, b->stream));
        NRH_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        NRH_HIP(hipEventElapsedTime(&ms, e0, e1));
        b->tuneMs[i] = ms / 3.0;
        if (ms / 3.0 < bestMs) {
            bestMs = ms / 3.0;
            bestK = cand[i];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tunedBlocksPerXcd = bestK;
    batch_set_sweep_blocks(b, bestK);
}

nrh_batch *batch_new(nrh_f3d *const *pairs, int n) {
    NRH_REQUIRE(pairs && n >= 1, NRH_ERR_ARG, "a batch needs at least one pair");
    std::unique_ptr<nrh_batch> b(new nrh_batch);
    b->P = n;
    b->pairs.assign(pairs, pairs + n);
    for (int q = 0; q < n; ++q) {
        NRH_REQUIRE(pairs[q] && !pairs[q]->velocity, NRH_ERR_ARG, "batched pairs are reg_f3d handles");
        for (int k = 0; k < q; ++k) NRH_REQUIRE(pairs[k] != pairs[q], NRH_ERR_ARG, "a handle appears twice in the batch");
        NRH_REQUIRE(pairs[q]->params.be_weight > 0.f, NRH_ERR_UNSUPPORTED, "batched pairs need a bending-energy term");
        NRH_REQUIRE(pairs[q]->device == pairs[0]->device, NRH_ERR_ARG, "batched pairs share one device");
    }
    b->device = pairs[0]->device;
    NRH_HIP(hipSetDevice(b->device));
    for (nrh_f3d *h : b->pairs) NRH_HIP(hipStreamSynchronize(h->stream));
    NRH_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    NRH_HIP(hipHostMalloc((void **)&b->published, (size_t)kBatchResults * n * sizeof(double),
                          hipHostMallocCoherent | hipHostMallocMapped));
    NRH_HIP(hipHostGetDevicePointer((void **)&b->publishedDev, b->published, 0));
    // every pair's work goes to the batch's stream while it belongs to the batch
    for (nrh_f3d *h : b->pairs) {
        b->ownStreams.push_back(h->stream);
        h->stream = b->stream;
    }
    try {
        batch_fused_setup(b.get());
        batch_ops_setup(b.get());
    } catch (...) {
        batch_free(b.release());  // gives the pairs their streams back
        throw;
    }
    return b.release();
}

void batch_free(nrh_batch *b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    for (size_t q = 0; q < b->pairs.size() && q < b->ownStreams.size(); ++q) b->pairs[q]->stream = b->ownStreams[q];
    if (b->ownsPairs)
        for (nrh_f3d *h : b->pairs) nrh_f3d_destroy(h);
    if (b->published) (void)hipHostFree(b->published);
    if (b->remaining) (void)hipHostFree(b->remaining);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // namespace nrh

using namespace nrh;

extern "C" {

int nrh_batch_create(nrh_f3d *const *pairs, int n, nrh_batch **out) {
    return guard([&] {
        NRH_REQUIRE(out, NRH_ERR_ARG, "null argument");
        *out = batch_new(pairs, n);
    });
}

int nrh_batch_destroy(nrh_batch *b) {
    return guard([&] { batch_free(b); });
}

int nrh_batch_objective(nrh_batch *b, double *values, double *nmi) {
    return guard([&] {
        NRH_REQUIRE(b, NRH_ERR_ARG, "null batch");
        batch_objective(b, nullptr);
        NRH_HIP(hipStreamSynchronize(b->stream));
        batch_touch(b, false);
        for (int q = 0; q < b->P; ++q) {
            if (values) values[q] = b->published[q * kBatchResults + 6];
            if (nmi) nmi[q] = b->published[q * kBatchResults + 0];
        }
    });
}

int nrh_batch_gradient(nrh_batch *b, double *max_lengths) {
    return guard([&] {
        NRH_REQUIRE(b, NRH_ERR_ARG, "null batch");
        batch_gradient(b, nullptr, max_lengths != nullptr);
        NRH_HIP(hipStreamSynchronize(b->stream));
        batch_touch(b, true);
        if (max_lengths)
            for (int q = 0; q < b->P; ++q) max_lengths[q] = b->published[q * kBatchResults + 7];
    });
}

int nrh_batch_set_sweep_blocks(nrh_batch *b, int blocks_per_xcd) {
    return guard([&] {
        NRH_REQUIRE(b && blocks_per_xcd >= 1, NRH_ERR_ARG, "invalid argument");
        batch_set_sweep_blocks(b, blocks_per_xcd);
    });
}

int nrh_batch_autotune(nrh_batch *b, int *blocks_per_xcd, double ms[4]) {
    return guard([&] {
        NRH_REQUIRE(b, NRH_ERR_ARG, "null batch");
        batch_autotune(b);
        NRH_HIP(hipStreamSynchronize(b->stream));
        batch_touch(b, false);
        if (blocks_per_xcd) *blocks_per_xcd = b->tunedBlocksPerXcd;
        if (ms)
            for (int i = 0; i < 4; ++i) ms[i] = b->tuneMs[i];
    });
}

}  // extern "C"
