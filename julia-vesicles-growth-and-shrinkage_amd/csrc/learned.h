// Learned preconditioner for the 2D psi BiCGSTAB: a 4-layer periodic 3x3
// convolutional network (2 -> 16 -> 16 -> 16 -> 1 channels, GELU), BF16 MFMA
// with FP32 accumulation for inference and training.  No reference counterpart.
#pragma once
#include <vector>

#include "common.h"

namespace sav {

typedef uint16_t bf16_t;

struct NetLayerDev {
  int cin, cout, kp, kpt;   // channels, padded K of the forward and of the transposed operand
  float* W;                 // master weights [cout][9*cin]  (k = tap*cin + c, tap = di*3 + dj)
  float* b;                 // [cout]
  float *mW, *vW, *mb, *vb; // Adam moments
  float* gW;                // gradient [cout][9*cin]
  float* gb;                // [cout]
  bf16_t* Wm;               // forward A operand  [16][kp]
  bf16_t* Wt;               // backward-data A operand [16][kpt]: Wt[c][tap'*cout + o] = W[o][(8-tap')*cin + c]
};

class Net {
 public:
  static constexpr int NL = 4;
  static constexpr int CH[NL + 1] = {2, 16, 16, 16, 1};
  static constexpr int NPARAM = 2 * 16 * 9 + 16 + 16 * 16 * 9 + 16 + 16 * 16 * 9 + 16 + 16 * 9 + 1;

  explicit Net(hipStream_t s);
  ~Net();

  // flat parameter vector: for each layer W[cout][cin][3][3] then b[cout]
  void set_params(const float* host);
  void get_params(float* host);

  // y[img][H][W] (fp32) = net(x), x = bf16 NHWC [img][H][W][2].  Activations are
  // kept for a following backward() when `keep` is set.
  void forward(const bf16_t* x, int nimg, int H, int W, float* y, bool keep);
  // Masked MSE against target (interior: border pixels excluded), gradients into gW/gb,
  // returns the loss (one device->host copy).
  double loss_backward(const bf16_t* x, const float* target, int nimg, int H, int W, int border);
  void adam_step(float lr, float beta1, float beta2, float eps);
  // gradient vector in the flat parameter layout (test hook)
  void get_grads(float* host);

  hipStream_t st;
  NetLayerDev L[NL];
  int t_adam = 0;

 private:
  void ensure(long long npix);
  void pack_operands();
  long long cap = 0;
  bf16_t* act[NL] = {};   // activations a_1..a_3 (16 ch), index 0 unused
  bf16_t* zb[NL] = {};    // pre-activations z_1..z_3
  bf16_t* g16a = nullptr; // gradient buffers (16 ch)
  bf16_t* g16b = nullptr;
  bf16_t* g1 = nullptr;   // last-layer gradient (1 ch)
  float* yb = nullptr;
  float* partials = nullptr;
  double* dloss = nullptr;
  std::vector<void*> allocs;
  template <class T>
  T* alloc(size_t n);
};

}  // namespace sav
