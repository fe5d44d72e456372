// guide_host.h -- host interface of the guiding network state (guiding.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "guide_mlp.h"
#include "guiding.h"

namespace lucid {
namespace guide {

struct State;

// Device state: FP32 master parameters with Adam moments, packed BF16 MFMA
// operands, a ring of `capacity` training samples and batch buffers for up
// to `batch_cap` samples.  nullptr when device memory runs out.
State* create(int capacity, int batch_cap);
void destroy(State* s);
// Fresh network (He-uniform, seeded), Adam state and sample ring cleared;
// the selection output starts at the guide fraction alpha0 everywhere.
void reset(State* s, uint32_t seed, float alpha0, hipStream_t st);
// Kernel view of the network for the scene box [bmin, bmax]: guide fraction
// alpha, or (learn_sel) the network's own per surface point.
GuideDev dev(const State* s, const float bmin[3], const float bmax[3], float alpha, bool learn_sel);
Sample* ring(State* s);
unsigned long long* ring_total(State* s);
int ring_capacity(const State* s);
// `steps` Adam steps on batches drawn uniformly from the ring.
void train(State* s, const GuideDev& g, int steps, int batch, float lr, hipStream_t st);
// One forward / backward pass over exactly `samples` with `params`, the
// selection gradient on: raw outputs (n x 40), parameter gradient (kParams)
// and loss.  For tests.
int debug_step(State* s, const float* params, const Sample* samples, int n, const float bmin[3], const float bmax[3],
               float* out, float* grad, float* loss, hipStream_t st);
// The rendering-side forward pass on explicit samples: raw outputs (n x 40).
int debug_infer(State* s, const float* params, const Sample* samples, int n, const float bmin[3], const float bmax[3],
                float* out, hipStream_t st);
int get_params(State* s, float* params, hipStream_t st);
float last_loss(State* s, hipStream_t st);
int set_params(State* s, const float* params, hipStream_t st);

}  // namespace guide
}  // namespace lucid
