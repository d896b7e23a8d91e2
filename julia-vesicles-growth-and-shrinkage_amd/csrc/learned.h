// Learned preconditioner for the 2D psi BiCGSTAB: a 4-layer periodic 3x3
// convolutional network (2 -> 16 -> 16 -> 16 -> 1 channels, GELU), BF16 MFMA
// with FP32 accumulation for inference and training.  No reference counterpart.
#pragma once
#include <vector>

#include "common.h"

namespace sav {

typedef uint16_t bf16_t;

// Training windows of the preconditioner are PATCH x PATCH pixels; the loss skips
// PATCH_BORDER pixels at each edge (the receptive-field radius of four 3x3 layers).
constexpr int PATCH = 64;
constexpr int PATCH_BORDER = 4;

__host__ __device__ __forceinline__ float bf2f(bf16_t h) {
  uint32_t u = ((uint32_t)h) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
// FP32 -> BF16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

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
  // inputs: mobility channel and a constant channel; outputs: 9 channels, of which the
  // preconditioner uses the first (the coefficient of its divergence-form stencil)
  static constexpr int CH[NL + 1] = {2, 16, 16, 16, 9};
  static constexpr int COUT = 9;
  static constexpr int NPARAM = 2 * 16 * 9 + 16 + 16 * 16 * 9 + 16 + 16 * 16 * 9 + 16 + 9 * 16 * 9 + 9;

  explicit Net(hipStream_t s);
  ~Net();

  // flat parameter vector: for each layer W[cout][cin][3][3] then b[cout]
  void set_params(const float* host);
  void get_params(float* host);
  // He-uniform hidden layers from a SplitMix64 stream, zero output layer (so the
  // untrained network adds nothing to the diagonal preconditioner), zero biases.
  void init_random(uint64_t seed);

  // y[img][H][W][COUT] (fp32) = net(x), x = bf16 NHWC [img][H][W][2].  Activations are
  // kept for a following backward pass when `keep` is set.
  void forward(const bf16_t* x, int nimg, int H, int W, float* y, bool keep);
  // Masked MSE of the raw outputs against target[img][H][W][COUT] (test hook); gradients
  // into gW/gb; returns the loss (one device->host copy).
  double loss_backward(const bf16_t* x, const float* target, int nimg, int H, int W, int border);
  // Masked MSE of the divergence-form stencil prediction -div(c grad v) (5-point, face
  // averages of the coefficient c = the first output channel; rx = dy/dx, ry = dx/dy;
  // periodic in each image) against target[img][H][W]; gradients into gW/gb; returns the loss.
  double loss_backward_stencil(const bf16_t* x, const float* v, const float* target, int nimg, int H, int W,
                               int border, double rx, double ry);
  void adam_step(float lr, float beta1, float beta2, float eps);
  // gradient vector in the flat parameter layout (test hook)
  void get_grads(float* host);

  hipStream_t st;
  NetLayerDev L[NL];
  int t_adam = 0;

 private:
  void ensure(long long npix);
  void pack_operands();
  double backward(const bf16_t* x, int nimg, int H, int W);
  long long cap = 0;
  bf16_t* act[NL] = {};   // activations a_1..a_3 (16 ch), index 0 unused
  bf16_t* zb[NL] = {};    // pre-activations z_1..z_3
  bf16_t* g16a = nullptr; // gradient buffers (16 ch)
  bf16_t* g16b = nullptr;
  bf16_t* glast = nullptr; // output-layer gradient (COUT ch)
  float* yb = nullptr;
  float* dres = nullptr;   // residuals of the divergence-form loss
  float* partials = nullptr;
  double* dloss = nullptr;
  int nb_loss = 0;
  std::vector<void*> allocs;
  template <class T>
  T* alloc(size_t n);
};

}  // namespace sav
