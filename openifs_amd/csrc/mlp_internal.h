// Internal state of a cloudsc_mlp_t, shared by the MLP and the surrogate translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace mlp {
typedef __bf16 bf16;
}  // namespace mlp

struct cloudsc_mlp {
  int nl = 0;                   // number of linear layers (hidden layers + 1)
  std::vector<int> din, dout;   // per layer
  int max_batch = 0;
  long step = 0;
  // per layer
  std::vector<float*> w, b, mw, vw, mb, vb, gw, gb;
  std::vector<mlp::bf16*> wb, wT;
  // per hidden layer l (0..nl-2): Z_l (fp32), A_l (mlp::bf16), A_l^T (mlp::bf16) -- A_l is the input of layer l+1
  std::vector<float*> z;
  std::vector<mlp::bf16*> a, aT;
  mlp::bf16 *g0 = nullptr, *g0T = nullptr, *g1 = nullptr, *g1T = nullptr;  // gradient ping-pong buffers
  float* y = nullptr;           // output [max_batch x n_out]
  float* loss = nullptr;        // device scalar
  float* ws = nullptr;          // partial products of split weight-gradient contractions
  size_t ws_elems = 0;
  std::vector<void*> allocs;
};

namespace mlp {
// Forward pass over M rows of x (bf16 [M x n_in]) into y (fp32 [M x n_out]); train keeps what backward needs.
hipError_t forward(cloudsc_mlp* m, const bf16* x, int M, float* y, bool train, hipStream_t s);
}  // namespace mlp
