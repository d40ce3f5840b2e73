/* ps_nn.h -- C-ABI of the kernels of the Stokes warm-start network (libpsamd.so).
 *
 * No counterpart in the reference (ParallelStencil.jl has no learned components): these
 * kernels serve a U-Net that predicts a starting point for the pseudo-transient Stokes
 * iteration of miniapps/Stokes2D.jl:109-125 from the viscosity field. They are plain
 * device-pointer entry points: BF16 storage (uint16 bit patterns), FP32 accumulation, NHWC
 * activations, weights of a 3x3 convolution stored [9*Cin][Cout] with row index
 * (ky*3 + kx)*Cin + ci. Every function returns 0 or a negative PS_E* code (ps_last_error()
 * gives the message); `stream` NULL is the library's stream 0. All launches are asynchronous
 * except ps_nn_h2d / ps_nn_d2h, which complete before returning.
 */
#ifndef PS_NN_H
#define PS_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GEMM epilogues */
#define PS_NN_EPI_BF16 0           /* C (bf16) = A B                              */
#define PS_NN_EPI_BF16_BIAS_RELU 1 /* C (bf16) = max(A B + bias[n], 0)            */
#define PS_NN_EPI_BF16_BIAS 2      /* C (bf16) = A B + bias[n]                    */
#define PS_NN_EPI_F32_ADD 3        /* C (fp32) += A B   (atomic; allows split-K)  */
#define PS_NN_EPI_F32 4            /* C (fp32) = A B                              */

/* C[m*ldc + n] = epi(sum_k A(m,k) B(k,n)) with A(m,k) = A[m*sam + k*sak] and
 * B(k,n) = B[k*sbk + n*sbn] (bf16); MFMA 16x16x32 bf16 with FP32 accumulation.
 * splitk > 1 splits K over workgroups (only with PS_NN_EPI_F32_ADD). */
int ps_nn_gemm(int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B, int64_t sbk, int64_t sbn, void* C,
               int64_t ldc, const float* bias, int epi, int splitk, void* stream);

/* ps_nn_gemm with A the 3x3 im2col of the NHWC input x (ps_nn_im2col's col, never stored):
 * transpose 0: A = col (M = B*H*W pixels, K = 9*C); transpose 1: A = col^T (M = 9*C,
 * K = B*H*W), the weight gradient's operand. Same products in the same order as
 * ps_nn_im2col followed by ps_nn_gemm, so the same results bit for bit. No counterpart in the
 * reference (the learned warm start has none); replaces the im2col buffer of the convolutions. */
int ps_nn_gemm_im2col(int transpose, int M, int N, int K, const void* x, int H, int W, int C, const void* B, int64_t sbk, int64_t sbn,
                      void* Cm, int64_t ldc, const float* bias, int epi, int splitk, void* stream);

/* 3x3, stride 1, zero padding 1, NHWC: col[(b*H*W + y*W + x)*9*C + (ky*3+kx)*C + c] */
int ps_nn_im2col(const void* x, int B, int H, int W, int C, void* col, void* stream);
/* weights of the data gradient of a 3x3 convolution with weights w (9*Cin x Cout, row
 * (ky*3+kx)*Cin + ci): taps reversed, wf[(t*Cout + co)*Cin + ci] =
 * w[((8-t)*Cin + ci)*Cout + co], so dx = ps_nn_gemm_im2col(0, ..., dy, ..., wf, ...) */
int ps_nn_conv_flip(const void* w, int Cin, int Cout, void* wf, void* stream);
/* adjoint of ps_nn_im2col: dx = sum of the column entries that read each input element */
int ps_nn_col2im(const void* dcol, int B, int H, int W, int C, void* dx, void* stream);
/* 2x2 average pooling of an HxW input (H, W even) and its adjoint */
int ps_nn_pool2(const void* x, int B, int H, int W, int C, void* y, void* stream);
int ps_nn_pool2_bwd(const void* dy, int B, int H, int W, int C, void* dx, void* stream);
/* nearest-neighbour upsampling of an HxW input to 2Hx2W and its adjoint */
int ps_nn_up2(const void* x, int B, int H, int W, int C, void* y, void* stream);
int ps_nn_up2_bwd(const void* dy, int B, int H, int W, int C, void* dx, void* stream);
/* channel concatenation of P pixels (a: Ca channels, b: Cb channels) and its adjoint */
int ps_nn_concat(const void* a, int Ca, const void* b, int Cb, int64_t P, void* y, void* stream);
int ps_nn_split(const void* dy, int Ca, int Cb, int64_t P, void* da, void* db, void* stream);
/* dy *= (y > 0) in place (ReLU backward from the ReLU output) */
int ps_nn_relu_bwd(void* dy, const void* y, int64_t n, void* stream);
/* db[n] += sum_m dy[m*N + n] (fp32) */
int ps_nn_bias_grad(const void* dy, int64_t M, int N, float* db, void* stream);
/* *loss += scale * sum mask*(pred - target)^2, dpred = 2*scale*mask*(pred - target); mask may be NULL */
int ps_nn_mse(const void* pred, const float* target, const float* mask, int64_t n, float scale, void* dpred, float* loss, void* stream);
/* Adam step on fp32 weights (c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)); w16 (optional) receives the bf16 copy */
int ps_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
               void* w16, void* stream);
int ps_nn_f32_to_bf16(const float* x, int64_t n, void* y, void* stream);
int ps_nn_bf16_to_f32(const void* x, int64_t n, float* y, void* stream);
int ps_nn_fill_f32(float* x, int64_t n, float v, void* stream);
/* y += x (bf16, summed in fp32) */
int ps_nn_add_bf16(void* y, const void* x, int64_t n, void* stream);

/* device buffers (zero-filled) and copies */
int ps_nn_alloc(int64_t bytes, void** out);
int ps_nn_free(void* p);
int ps_nn_h2d(void* dst, const void* src, int64_t bytes, void* stream);
int ps_nn_d2h(void* dst, const void* src, int64_t bytes, void* stream);
int ps_nn_d2d(void* dst, const void* src, int64_t bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
