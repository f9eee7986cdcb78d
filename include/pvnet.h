/* pvnet.h — C ABI of the warm-start network kernels (libpvnet.so, gfx950).
 *
 * The reference has no counterpart: its solver always starts from rest (Simulation.cpp:13-15 sets
 * every population to its weight before the time loop).  These entry points are the compute of a
 * small 3D U-Net that maps voxel geometry plus inlet parameters to an initial (u, rho) field; the
 * solver of pvlbm.h runs on from that field (pvlbm_init_fields) and decides the result.
 *
 * Tensors are batches of boxes, channels-last: element (b, z, y, x, c) of a tensor with row
 * length ld sits at ((((b*d + z)*h + y)*w + x)*ld + c).  Activations and activation gradients are
 * bfloat16 (passed as uint16_t bit patterns); accumulation is FP32 in the MFMA units; master
 * weights, weight gradients and optimizer moments are FP32.  Channel counts and channel offsets
 * must be multiples of 8 (16-byte rows of 8 bf16), and the voxel count nb*d*h*w a multiple of 32.
 *
 * Every call is asynchronous on the given HIP stream (NULL = the null stream) and takes device
 * pointers only; none allocates.  Return codes are those of pvlbm.h (0 = OK, 1 = bad argument,
 * 2 = HIP error); pvnet_last_error() gives the text of the last failure on the calling thread.
 */
#ifndef PVNET_H
#define PVNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t nb, d, h, w; /* batch of nb boxes of d x h x w voxels (z, y, x; x fastest) */
} pvnet_shape;

enum { PVNET_RELU = 1, PVNET_F32OUT = 2, PVNET_GATHER = 4 };

/* y[m][n] = act(bias[n] + sum_{tap, ci} x[m + off(tap)][ci] * w[n][tap*cin + ci])
 * ntaps = 27: 3x3x3 stencil with zero padding, tap = (dz+1)*9 + (dy+1)*3 + (dx+1);
 * ntaps = 1: pointwise.  w: bf16 [cout][kpad], kpad = roundup(ntaps*cin, 32), zero beyond
 * ntaps*cin.  If relu_of is not NULL the result is multiplied by (relu_of[m][n] > 0) (the
 * backward of a ReLU whose output is relu_of).  flags: PVNET_RELU applies max(., 0);
 * PVNET_F32OUT writes float instead of bf16.  y may be a channel slice of a wider tensor.
 * 3x3x3 convolutions with 16, 32, 48 or 64 outputs (or multiples of 48 or 64) on grids of whole
 * 16 x 4 x 4 bricks (x, y, z; 16 x 8 x 4 where the halo fits) stage each brick and its halo in
 * LDS once; PVNET_GATHER keeps them on the per-tap gathering kernel (the same result bit for bit,
 * for comparison).  Pointwise convolutions with cin <= 32 and cout <= 64 take their matrix-core
 * operands straight from global memory (no LDS); PVNET_GATHER keeps them on the tiled kernel
 * (again the same result bit for bit). */
typedef struct {
    pvnet_shape s;
    int32_t ntaps, cin, cout, kpad;
    const uint16_t* x;
    int32_t ldx;
    int32_t flags;
    const uint16_t* w;
    const float* bias;
    void* y;
    int32_t ldy;
    int32_t ldr;
    const uint16_t* relu_of;
} pvnet_conv;

int pvnet_conv_fwd(const pvnet_conv* c, void* stream);

/* dw[n][k] = sum_m dy[m][n] * x[m + off(tap(k))][ci(k)]  for k < ntaps*cin (row length ntaps*cin)
 * db[n]    = sum_m dy[m][n]                               (if db is not NULL; the 16-output brick
 *                                                          kernels sum it in the same pass)
 * ws: FP32 scratch of at least pvnet_wgrad_ws_floats(...) floats for the split-K partials; the
 * sums are reduced in a fixed order, so results are deterministic.  16- and 32-output 3x3x3 layers
 * on grids of whole 8 x 8 x 4 bricks reduce brick by brick with the input staged in LDS;
 * flags = PVNET_GATHER keeps them on the per-tap gathering kernel (the same sums in another
 * order, for comparison).  flags = 0 otherwise. */
typedef struct {
    pvnet_shape s;
    int32_t ntaps, cin, cout;
    const uint16_t* x;
    int32_t ldx;
    int32_t lddy;
    const uint16_t* dy;
    float* ws;
    float* dw;
    float* db;
    int32_t flags;
} pvnet_wgrad;

int64_t pvnet_wgrad_ws_floats(const pvnet_shape* s, int32_t ntaps, int32_t cin, int32_t cout);
int pvnet_conv_wgrad(const pvnet_wgrad* g, void* stream);

/* Packs FP32 master weights w32[cout][ntaps*cin] into the bf16 forward layout wt[cout][kpad]
 * and the bf16 input-gradient layout wd[cin][kpad_d] (kpad_d = roundup(ntaps*cout, 32)), whose
 * conv with pvnet_conv_fwd (cin and cout swapped) is the transposed convolution:
 * wd[ci][tap'*cout + co] = w32[co][(ntaps-1-tap')*cin + ci]. */
int pvnet_pack_weights(const float* w32, int32_t ntaps, int32_t cin, int32_t cout, uint16_t* wt,
                       uint16_t* wd, void* stream);

/* 2x2x2 average pooling: y (shape s/2) from x (shape s); backward adds dy/8 to each child of
 * dx (accumulate = 1) or overwrites it (accumulate = 0).  All three sizes of s must be even. */
int pvnet_pool2_fwd(const pvnet_shape* s, int32_t c, const uint16_t* x, int32_t ldx, uint16_t* y,
                    int32_t ldy, void* stream);
int pvnet_pool2_bwd(const pvnet_shape* s, int32_t c, const uint16_t* dy, int32_t lddy, uint16_t* dx,
                    int32_t lddx, int32_t accumulate, void* stream);
/* The accumulating backward followed by the backward of a ReLU whose output is relu_of, in one
 * pass: dx = (relu_of > 0) ? dx + dy/8 : 0 (the same values as pvnet_pool2_bwd with accumulate = 1
 * and then pvnet_relu_bwd on dx). */
int pvnet_pool2_bwd_relu(const pvnet_shape* s, int32_t c, const uint16_t* dy, int32_t lddy, uint16_t* dx,
                         int32_t lddx, const uint16_t* relu_of, int32_t ldr, void* stream);

/* nearest 2x upsampling: y (shape s) from x (shape s/2); backward sums the 8 children. */
int pvnet_up2_fwd(const pvnet_shape* s, int32_t c, const uint16_t* x, int32_t ldx, uint16_t* y,
                  int32_t ldy, void* stream);
int pvnet_up2_bwd(const pvnet_shape* s, int32_t c, const uint16_t* dy, int32_t lddy, uint16_t* dx,
                  int32_t lddx, void* stream);

/* dx[m][c] = (a[m][c] > 0) ? dx[m][c] : 0, in place (the backward of a ReLU with output a). */
int pvnet_relu_bwd(int64_t m, int32_t c, const uint16_t* a, int32_t lda, uint16_t* dx, int32_t lddx,
                   void* stream);

/* Euclidean distance (in voxels) from each voxel to the nearest solid voxel (code 0) of its box,
 * exact: the separable squared-distance transform of Saito and Toriwaki (1994), one pass per
 * axis; capped at 255 (a box without solid voxels).  Box sizes up to PVNET_EDT_MAX per axis. */
#define PVNET_EDT_MAX 512
int pvnet_wall_distance(const pvnet_shape* s, const uint8_t* code, float* dist, void* stream);

/* Input features (8 bf16 channels) from a geometry code per voxel (0 solid, 1 fluid, 2 inlet
 * face, 3 outlet face), its wall distance (pvnet_wall_distance) and per-box parameters
 * prm[b] = {u_mean, tau, inlet_area} (inlet_area: number of inlet-face voxels):
 *   c0 fluid, c1 inlet, c2 outlet, c3 fluid*(tau-0.5)/0.4, c4 fluid*u_mean/0.05,
 *   c5 fluid*(x/(w-1) - 0.5), c6 fluid*dist/16, c7 fluid*inlet_area/(pi*16^2). */
int pvnet_features(const pvnet_shape* s, const uint8_t* code, const float* prm, const float* dist,
                   uint16_t* x, void* stream);

/* Masked mean-squared error over fluid voxels (code > 0) and the first nc channels of y (FP32,
 * row length ldy) against t (bf16 [m][nc]) with channel weights wc[nc]:
 *   L = sum_{fluid m, c} wc[c] (y - t)^2 / (n_fluid * nc)
 * dy (bf16, row length lddy) receives dL/dy on channels [0, nc) and zero on [nc, lddy).
 * loss[0] = L, loss[1] = n_fluid (FP64 sums of FP32 terms in a fixed order); part: scratch of
 * pvnet_loss_ws_doubles(m) doubles. */
int64_t pvnet_loss_ws_doubles(int64_t m);
int pvnet_mse_loss(int64_t m, int32_t nc, const float* y, int32_t ldy, const uint16_t* t,
                   const uint8_t* code, const float* wc, uint16_t* dy, int32_t lddy, double* part,
                   double* loss, void* stream);

/* Adam (Kingma & Ba 2015) on n FP32 parameters, step t >= 1, bias-corrected:
 *   m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;  p -= lr * (m/(1-b1^t)) / (sqrt(v/(1-b2^t)) + eps) */
int pvnet_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
               float eps, int32_t t, void* stream);

const char* pvnet_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
