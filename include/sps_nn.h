/*
 * sps_nn.h - C ABI of the BF16 MFMA convolution kernels behind the learned
 * surrogate of the SolverFT step.
 *
 * No counterpart in the reference: SymPhas has no learned components.  The
 * surrogate is a convolutional neural operator that stands in for k-step
 * advances of the finite-difference solver (its fallback and referee); these
 * entry points are its forward and backward building blocks, called by the
 * host-side training and inference code with device pointers.
 *
 * Tensors are NHWC (channels fastest), BF16 stored as uint16 bit patterns,
 * with periodic (circular) padding of the 3x3 neighbourhood, matching the
 * periodic fields of the solver.  Weights of a 3x3 convolution are laid out
 * [tap = 3*r + s][out][in] with
 *     y[n,h,w,k] = bias[k] + sum_{r,s,c} x[n, h+r-1, w+s-1, c] * W[3r+s][k][c]
 * (indices taken modulo H and W).  Products are exact BF16 x BF16 and sums
 * accumulate in FP32 inside v_mfma_f32_16x16x32_bf16.
 *
 * Every function returns an sps_status (include/sps_fd.h) and enqueues its
 * work on the given hipStream_t (NULL = the null stream); sps_fd_last_error()
 * holds the message of the last failure.
 */
#ifndef SPS_NN_H
#define SPS_NN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SPS_NN_ACT_NONE = 0, SPS_NN_ACT_SILU = 1 };
enum { SPS_NN_EPI_PLAIN = 0, SPS_NN_EPI_UPSAMPLE_ADD = 1 };

/* 3x3 convolution, cin % 32 == 0, cout % 32 == 0, n*h*w % 64 == 0.
 * epilogue PLAIN: pre = acc + bias (optional output), y = act(pre), both
 *   [n][h][w][cout].
 * epilogue UPSAMPLE_ADD: nearest 2x upsampling with a skip connection,
 *   pre[n][2h+a][2w+b][k] = acc[n][h][w][k] + bias[k] + skip[n][2h+a][2w+b][k],
 *   y = act(pre); skip, pre and y are [n][2h][2w][cout].
 * bias and pre may be NULL. */
int sps_nn_conv3x3(const void* x, int n, int h, int w, int cin,
                   const void* weight, const float* bias, int cout, int act,
                   int epilogue, const void* skip, void* y, void* pre,
                   void* stream);

/* Weight gradients are deterministic: each workgroup reduces its share of
 * the pixels in a fixed order into its own partial gradient in a caller-owned
 * FP32 workspace; a fixed tree then adds the partials to dweight (runs of
 * 32 partials summed in order, in place in the workspace, until at most 32
 * remain, which are added to dweight in order).  The same inputs give the
 * same bits on every call.  The workspace is scratch: its contents are
 * consumed.
 * sps_nn_wgrad_workspace gives the workspace size in floats for a layer
 * (kind SPS_NN_WGRAD_CONV with cin and cout, SPS_NN_WGRAD_IN1 with cout,
 * SPS_NN_WGRAD_OUT1 with cin); 0 for invalid sizes. */
#define SPS_NN_WGRAD_CONV 0
#define SPS_NN_WGRAD_IN1 1
#define SPS_NN_WGRAD_OUT1 2
size_t sps_nn_wgrad_workspace(int kind, int n, int h, int w, int cin,
                              int cout);

/* Weight gradient of sps_nn_conv3x3 (PLAIN epilogue):
 *   dweight[3r+s][k][c] += sum_{n,h,w} x[n,h+r-1,w+s-1,c] * dpre[n,h,w,k]
 * accumulated in FP32 into dweight.  n*h*w % 32 == 0. */
int sps_nn_conv3x3_wgrad(const void* x, const void* dpre, int n, int h, int w,
                         int cin, int cout, float* dweight, float* workspace,
                         size_t ws_floats, void* stream);

/* First layer, one FP32 input channel:
 *   pre[n,h,w,k] = bias[k] + sum_t x[n, shift_t(h,w)] * weight[t][k],
 *   y = act(pre) (BF16), weight [9][cout] FP32, cout % 8 == 0, cout <= 64. */
int sps_nn_conv3x3_in1(const float* x, int n, int h, int w,
                       const float* weight, const float* bias, int cout,
                       int act, void* y, void* pre, void* stream);
/* dweight[t][k] += sum x[shift_t(p)] * dpre[p][k] (FP32 accumulate). */
int sps_nn_conv3x3_in1_wgrad(const float* x, const void* dpre, int n, int h,
                             int w, int cout, float* dweight, float* workspace,
                             size_t ws_floats, void* stream);

/* Last layer, one FP32 output channel:
 *   y[n,h,w] = bias[0] + sum_{t,c} x[n, shift_t(h,w), c] * weight[t][c],
 *   weight [9][cin] FP32, cin % 8 == 0, cin <= 64. */
int sps_nn_conv3x3_out1(const void* x, int n, int h, int w, int cin,
                        const float* weight, const float* bias, float* y,
                        void* stream);
/* Backward of the last layer: dx (BF16, [n][h][w][cin]) and
 * dweight[t][c] += sum_p x[shift_t(p)][c] * dy[p] (FP32). */
int sps_nn_conv3x3_out1_bwd(const void* x, const float* dy, int n, int h,
                            int w, int cin, const float* weight, void* dx,
                            float* dweight, float* workspace, size_t ws_floats,
                            void* stream);

/* dx = dy * silu'(pre), all BF16, count elements. */
int sps_nn_silu_bwd(const void* dy, const void* pre, void* dx, size_t count,
                    void* stream);
/* 2x2 pooling of [n][h][w][c] into [n][h/2][w/2][c]: mode 0 average, 1 sum. */
int sps_nn_pool2(const void* x, int n, int h, int w, int c, int mode, void* y,
                 void* stream);
/* Broadcast [n][h][w][c] to [n][2h][2w][c] times scale (the adjoint of
 * average pooling with scale 1/4). */
int sps_nn_unpool2(const void* x, int n, int h, int w, int c, float scale,
                   void* y, void* stream);

/* y[s][i] = x[s][i] - mean_i x[s][i] for each of n samples of len FP32
 * values: the mass-conserving projection of the surrogate's increment (it is
 * its own adjoint, so the backward pass applies it to the gradient).  The
 * mean is an FP64 sum in a fixed order, one workgroup per sample; x and y may
 * alias. */
int sps_nn_center(const float* x, float* y, int n, int64_t len, void* stream);

#ifdef __cplusplus
}
#endif

#endif
