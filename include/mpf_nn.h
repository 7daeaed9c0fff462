/* mpf_nn.h -- device kernels of the learned surrogate (libmpf.so).
 *
 * No counterpart in the reference: the reference has no learned model.
 * These entry points serve the slot-equivariant 3D U-Net that proposes
 * 16-step advances of the phase field (mpf/surrogate.py).  All pointers
 * are device pointers; stream may be NULL (default stream); calls are
 * asynchronous like kernel launches.  Return 0, -1 for bad arguments, -2
 * for a launch error.
 *
 * Tensors are channels-last [B][X][Y][Z][C] (z fastest), bf16 unless
 * stated.  dims = {B, X, Y, Z}; periodic = {px, py, pz} (1: periodic axis,
 * 0: zero padding).  Convolution weights are fp32 masters [ntap][cout][cin]
 * converted to bf16 by mpf_nn_weights_bf16; ntap is 27 (3x3x3) or 1
 * (1x1x1); cin and cout are multiples of 32.
 */
#ifndef MPF_NN_H
#define MPF_NN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y = relu?(conv(x, w) + bias): y bf16 and/or yf fp32 (either may be NULL). */
int mpf_nn_conv_fwd(const void *x, const void *w, const float *bias, void *y, float *yf, const int *dims, int cin,
                    int cout, int ntap, const int *periodic, int relu, void *stream);
/* dw[tap][co][ci] += sum_v dy[v][co] * x[v + tap][ci] (fp32 dw, accumulated). */
int mpf_nn_conv_bwd_w(const void *x, const void *dy, float *dw, const int *dims, int cin, int cout, int ntap,
                      const int *periodic, void *stream);
/* bf16 weights from fp32 masters [ntap][cout][cin]; flip 1 writes the
 * backward-data weights [ntap][cin][cout] with the taps reversed. */
int mpf_nn_weights_bf16(const float *w, void *out, int ntap, int cout, int cin, int flip, void *stream);
/* dy = 0 where y <= 0 (in place). */
int mpf_nn_relu_bwd(void *dy, const void *y, size_t n, void *stream);
/* db[c] += sum_v dy[v][c] (fp32). */
int mpf_nn_bias_grad(const void *dy, float *db, int nvox, int c, void *stream);
/* 2x average pooling of x (dims = the input's) into y; backward: the
 * adjoint from the pooled gradient x (dims = the fine grid's) into y. */
int mpf_nn_pool(const void *x, void *y, const int *dims, int c, int backward, void *stream);
/* 2x nearest upsampling of x (dims = the coarse grid's) into y; backward:
 * the adjoint from the fine gradient x into the coarse y (dims = coarse). */
int mpf_nn_upsample(const void *x, void *y, const int *dims, int c, int backward, void *stream);
/* dst[v][dst_off + j] (=, or += with accumulate) src[v][src_off + j], j < nch. */
int mpf_nn_copy_channels(const void *src, int src_c, int src_off, void *dst, int dst_c, int dst_off, int nch,
                         int nvox, int accumulate, void *stream);
/* Adam step t (>= 1) on fp32 parameters. */
int mpf_nn_adam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                int t, void *stream);

/* Slot features of the band [z0, z0 + zb) of a dense float32
 * phi[n_phase][nx][ny][nz]: a cell's slots are the grains (id >= 1) with
 * phi > 0 at the cell or a face neighbour (x, y periodic, z clamped),
 * ascending id, at most k.  sid[k][nv] (-1 empty), feat[k][nv][32] =
 * (phi_g, orient[g][0..2], 0...), aux[nv][32] = (phi_liquid, tfeat[z], 1,
 * 0...), nv = nx*ny*zb; *overflow += cells with more than k grains. */
int mpf_nn_slot_features(const float *phi, int n_phase, int nx, int ny, int nz, int z0, int zb, int k,
                         const float *orient, const float *tfeat, int32_t *sid, void *feat, void *aux, int *overflow,
                         void *stream);
/* h[v] = sum over occupied slots of e[s][v] (c channels); backward: the
 * adjoint de[s][v] (=, or += with accumulate) dh[v] for occupied slots
 * (e is then dh and h is de). */
int mpf_nn_slot_pool(const void *e, const int32_t *sid, void *h, int nv, int k, int c, int backward, int accumulate,
                     void *stream);
/* hin[s][v] = (f[v][0..cf), e[s][v][0..ce)); backward: dF = sum over s,
 * de = the e part (hin is then the gradient, f and e the outputs). */
int mpf_nn_head_in(const void *f, int cf, const void *e, int ce, void *hin, int nv, int k, int backward,
                   void *stream);
/* tgt[s][v] = scale * (phi1 - phi0) at slot s's grain, mask 1; empty slots 0. */
int mpf_nn_slot_targets(const float *phi0, const float *phi1, int nx, int ny, int nz, int z0, int zb, int k,
                        const int32_t *sid, float scale, float *tgt, float *mask, void *stream);
/* Masked squared error of out[i*stride] against tgt[i]: *loss += sum,
 * dout[i*stride] = 2 (out - tgt) mask * inv_count, other channels 0. */
int mpf_nn_mse(const float *out, int stride, const float *tgt, const float *mask, size_t n, float inv_count,
               void *dout, float *loss, void *stream);
/* Band update of a dense float32 phi in place: slot grains += out/scale
 * clipped to [0, 1], liquid = 1 - sum of solids, solids above 1 rescaled. */
int mpf_nn_slot_apply(float *phi, int n_phase, int nx, int ny, int nz, int z0, int zb, int k, const int32_t *sid,
                      const float *out, int stride, float inv_scale, void *stream);
/* *out = max(*out, max |a - b| over the band) as float bits. */
int mpf_nn_band_maxdiff(const float *a, const float *b, int n_phase, int nx, int ny, int nz, int z0, int zb,
                        unsigned *out, void *stream);
/* *out = max(*out, max over the band of |(c - b) - (b - a) * inv_h|) as float
 * bits: the exact one-step change c - b at a proposal b against the
 * proposal's mean rate from a over 1/inv_h steps. */
int mpf_nn_band_residual(const float *a, const float *b, const float *c, int n_phase, int nx, int ny, int nz,
                         int z0, int zb, float inv_h, unsigned *out, void *stream);
int mpf_nn_to_bf16(const float *x, void *y, size_t n, void *stream);
int mpf_nn_fill(float *x, float v, size_t n, void *stream);
/* *out = sum of x[0, n) (float atomics per block: exact for integer-valued
 * sums below 2^24, e.g. the count of a 0/1 mask). */
int mpf_nn_sum(const float *x, size_t n, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
