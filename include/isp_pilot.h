/*
 * isp_pilot.h - C ABI of the learned pilot kernels in libisp_bm3d.so.
 *
 * A small residual CNN (3x3 convolutions, 48 channels, ReLU) can stand in for
 * BM3D_Basic's estimate (source/BM3D.cpp:339-385) as the guide of the Wiener
 * step BM3D_Final (source/BM3D.cpp:411-457); these are its building blocks.
 * None of them has a counterpart in ISP_MW.
 *
 * Activations are BF16 (raw 16-bit patterns) in NHWC layout [n][h][w][c] with
 * c a multiple of 16; arithmetic accumulates in FP32 on the matrix cores.
 * Every pointer is device memory; work is enqueued on `stream` (a
 * hipStream_t, NULL for the null stream) and nothing synchronises.
 *
 * Weights: FP32 master copy [cout][cin][9], tap = ky*3 + kx, cross-correlation
 *   out[y][x][co] = b[co] + sum w[co][ci][ky*3+kx] * in[y+ky-1][x+kx-1][ci]
 * with zeros outside the image.  isp_pilot_pack writes the BF16 forms the
 * convolution reads: wf [cout_pad][kpad(cin_pad)] and wb [cin_pad][kpad(cout_pad)]
 * (flipped and transposed, for the input gradient).
 */
#ifndef ISP_PILOT_H_
#define ISP_PILOT_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISP_PILOT_OK 0
#define ISP_PILOT_E_ARG (-1)
#define ISP_PILOT_E_UNSUPPORTED (-2)
#define ISP_PILOT_E_HIP (-4)

/* K extent of a packed weight row: 9 * cin_pad rounded up to 32 */
int isp_pilot_kpad(int cin_pad);

/* One 3x3 convolution.  (cin_pad, cout_pad) in {(16,48), (48,48), (48,16)}.
 * mask (may be NULL): input values whose mask value is not > 0 read as 0 (the
 * ReLU derivative of a post-ReLU activation, for the input gradient).
 * bias may be NULL; relu != 0 applies max(., 0); out_f32 (may be NULL) gets
 * the FP32 results too. */
int isp_pilot_conv3x3(const void *in, const void *mask, const void *weights, const float *bias, void *out,
                      float *out_f32, int n, int h, int w, int cin_pad, int cout_pad, int relu, void *stream);

/* The last layer at inference, fused with the output stage: for a
 * (cin_pad, cout_pad) = (48, 16) convolution whose first `channels` outputs
 * are the residuals of the image's planes, est[i][c][y][x] = noisy[i][c][y][x]
 * - (conv + bias)[c] (noisy and est FP32 [n][channels][h][w]), with the same
 * values as isp_pilot_conv3x3 into out_f32 followed by
 * isp_pilot_output_planes. */
int isp_pilot_conv3x3_residual(const void *in, const void *weights, const float *bias, const float *noisy,
                               float *est, int channels, int n, int h, int w, int cin_pad, int cout_pad,
                               void *stream);

/* Weight and bias gradients of a convolution with real channel counts
 * cin <= cin_pad, cout <= cout_pad (the same pads as isp_pilot_conv3x3), on
 * the matrix cores: dw [cout][cin][9], db [cout] (may be NULL), with
 * dz = dout * (mask > 0) (mask may be NULL).  partial: scratch of
 * isp_pilot_wgrad_partial_floats floats.  The per-workgroup sums are added
 * in a fixed order (run to run identical). */
size_t isp_pilot_wgrad_partial_floats(int n, int h, int w, int cin, int cout);
int isp_pilot_conv3x3_wgrad(const void *in, const void *dout, const void *mask, float *partial, float *dw, float *db,
                            int n, int h, int w, int cin_pad, int cout_pad, int cin, int cout, void *stream);

int isp_pilot_pack(const float *w, void *wf, void *wb, int cin, int cout, int cin_pad, int cout_pad, void *stream);

/* Adam (bias-corrected) on n parameters at step >= 1 */
int isp_pilot_adam(float *p, const float *g, float *m, float *v, long n, float lr, float beta1, float beta2,
                   float eps, int step, void *stream);

/* Network input [n][h][w][16] of n images of `channels` planes each (noisy:
 * [n][channels][h][w], 1 <= channels <= 15): channel c < channels the
 * noisy plane c plus offset[i * channels + c] (offset may be NULL: 0; chroma
 * planes centred on 0 pass 0.5), channel `channels` sigma[i]/255 (one sigma
 * per image, 0..255 units), the rest 0.  The residual output makes the
 * offsets cancel: out = noisy - r. */
int isp_pilot_input_planes(const float *noisy, int channels, const float *sigma, const float *offset, void *x, int n,
                           int h, int w, void *stream);
/* isp_pilot_input_planes with one plane per image */
int isp_pilot_input(const float *noisy, const float *sigma, const float *offset, void *x, int n, int h, int w,
                    void *stream);

/* out = noisy - residual per plane (plane c of image i: residual channel c of
 * res_f32 [n][h][w][16]; noisy, out, target [n][channels][h][w]).  With a
 * target: *loss += mean over every plane of (out - target)^2 and dres
 * [n][h][w][16] (bf16) gets d(loss)/d(residual) in channels 0..channels-1.
 * out may be NULL. */
int isp_pilot_output_planes(const float *noisy, const float *res_f32, float *out, const float *target, void *dres,
                            float *loss, int channels, int n, int h, int w, void *stream);
/* isp_pilot_output_planes with one plane per image */
int isp_pilot_output(const float *noisy, const float *res_f32, float *out, const float *target, void *dres,
                     float *loss, int n, int h, int w, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ISP_PILOT_H_ */
