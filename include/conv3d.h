/*
 * jaxphys_amd -- C ABI of the BF16-MFMA 3D convolution library
 * (libconv3d_hip.so) used by the learned dispersion closure.
 *
 * No reference counterpart: the reference (sushaan-k/jaxphys) has no neural
 * network or closure code.  These entry points implement the 3x3x3, stride
 * 1, zero-padded ("same") convolution of a small channels-last network as
 * implicit GEMMs on gfx950 matrix cores (v_mfma_f32_16x16x32_bf16) with
 * float32 accumulation, for inference and for training (input and weight
 * gradients).
 *
 * Layouts (all device pointers, caller-owned):
 *  activations: bf16, channels last, [nz][ny][nx][C] with C = the padded
 *               channel count (a multiple of 8);
 *  weights:     bf16, [cout_pad][kpad] with k = tap*cin_pad + c,
 *               tap = (dz+1)*9 + (dy+1)*3 + (dx+1), kpad = 27*cin_pad
 *               rounded up to a multiple of 32 (the tail is zero);
 *               cout_pad a multiple of 16;
 *  bias:        float32 [cout] (may be NULL);
 *  weight grad: float32 [cout_pad][kpad].
 * Return codes as in yee.h: 0 ok, -1 bad argument, -2 HIP error.
 */
#ifndef JAXPHYS_AMD_CONV3D_H
#define JAXPHYS_AMD_CONV3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* epilogue of conv3d_fwd */
#define CONV_EPI_GELU 0      /* z = acc + bias (bf16 to z_out if given); y = gelu(z) bf16 */
#define CONV_EPI_LINEAR_F32 1 /* y = acc + bias as float32, channel stride y_cpad (channels
                                 cout..y_cpad-1 written as zero when y_cpad is a multiple of 4) */
#define CONV_EPI_GELU_BWD 2  /* y = acc * gelu'(aux) bf16 (aux: bf16 pre-activation) */
#define CONV_EPI_LINEAR_BF16 3 /* y = acc + bias as bf16 */

typedef struct ConvDesc {
  int32_t nx, ny, nz;
  int32_t cin_pad;    /* channels of x (multiple of 8) */
  int32_t cout;       /* real output channels */
  int32_t cout_pad;   /* multiple of 16 */
  int32_t y_cpad;     /* channel stride of the output tensor (>= cout) */
} ConvDesc;

/* y = epilogue(conv(x, w) + bias) */
int conv3d_fwd(const ConvDesc* d, const void* x, const void* w, const float* bias, int32_t epi,
               void* z_out, const void* aux, void* y, void* stream);

/* dw[co][k] = sum_p dy[p][co] * x[p + off(tap(k))][c(k)]; db[co] = sum_p dy[p][co].
 * dy: bf16 channels last with channel stride dy_cpad (>= cout, a multiple
 * of 8); x and dy 16-byte aligned.  `work` is a float32 device buffer of
 * conv3d_wgrad_work_floats(d) floats (per-block partial sums of dw and db,
 * then a fixed-order reduction: deterministic); dw, db are overwritten.  */
int64_t conv3d_wgrad_work_floats(const ConvDesc* d);
int conv3d_wgrad(const ConvDesc* d, const void* x, const void* dy, int32_t dy_cpad, float* work,
                 float* dw, float* db, void* stream);

/* Weight transform for the input gradient: wt[ci][tap'*cout_pad_in + co] =
 * w[co][(26-tap')*cin_pad + ci]  (spatially flipped, channels swapped), so
 * that dx = conv3d_fwd(dy, wt).  wt is [cin_pad rounded to 16][kpad_t] with
 * kpad_t = 27*cout_pad_in rounded up to 32, cout_pad_in = dy channel stride
 * rounded up to 8.                                                        */
int conv3d_flip_weights(const ConvDesc* d, const void* w, int32_t dy_cpad, void* wt, void* stream);

/* Pack the six Yee fields and eps_r (float32, x fastest, (nz,ny,nx)) into a
 * bf16 channels-last tensor with 8 channels: c = field / scale[c] for
 * c < 6 (with CONV_PACK_ASINH in `flags`: asinh(field / scale[c]), which
 * keeps the values next to a point source, thousands of times the RMS, in
 * the range of the rest), eps channel (eps - eps_shift) / eps_scale,
 * channel 7 = 0.                                                          */
#define CONV_PACK_ASINH 1
int conv3d_pack_fields(int32_t nx, int32_t ny, int32_t nz, float* const* fields, const float* eps,
                       const float* scale, float eps_shift, float eps_scale, int32_t flags,
                       void* out, void* stream);
/* conv3d_pack_fields over the box (x0, y0, z0) + [0, bnx) x [0, bny) x
 * [0, bnz) of the grid, out in box order (bnz, bny, bnx, 8): the closure
 * net's input on the cells it is evaluated on.                            */
int conv3d_pack_fields_box(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                           const float* eps, const float* scale, float eps_shift, float eps_scale,
                           int32_t flags, int32_t x0, int32_t y0, int32_t z0, int32_t bnx,
                           int32_t bny, int32_t bnz, void* out, void* stream);
/* fields[c] += gain * scale[c] * y[..][c], y float32 channels last [..][ystride],
 * for the cells at least `margin` cells from every face; the others (the
 * absorbing layers, where the closure is not trained) are left unchanged. */
int conv3d_apply_correction(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                            const float* y, int32_t ystride, const float* scale, float gain,
                            int32_t margin, void* stream);

/* The closure's linear far-field operator (float32; no reference
 * counterpart): y[i][co] += sum over the 27 taps t = ((dz+1)*3 + dy+1)*3 +
 * dx+1 and the six fields ci of
 *   (w[t][ci][0][co] + (eps[i] - 1) w[t][ci][1][co]) * f_ci(i + (dx,dy,dz)) / scale[ci],
 * taps outside the grid contributing 0.  fields: 6 device pointers
 * (float32, x fastest); eps (per cell), scale (6), w (27*6*2*6 floats) and
 * y (float32 channels last [..][ystride]) on the device.                  */
int conv3d_stencil_f32(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                       const float* eps, const float* scale, const float* w, float* y,
                       int32_t ystride, void* stream);
/* conv3d_stencil_f32 with y[i][co] = (not +=) the operator's value, for
 * the cells of box (x0, y0, z0, bnx, bny, bnz; NULL: the whole grid) only
 * (y indexed on the whole grid; the other cells untouched).               */
int conv3d_stencil_f32_set(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                           const float* eps, const float* scale, const float* w, float* y,
                           int32_t ystride, const int32_t* box, void* stream);

/* One closure correction in one pass (no reference counterpart; the same
 * float operations as y = net_gain * (yb - y0b), y += s, y[mask] = 0 and
 * conv3d_apply_correction): for the cells at least `margin` from every
 * face,
 *   v[c] = net_gain * (yb[j][c] - y0b[j][c])   (yb, y0b: the net's output
 *          and its zero-state response, float32 [..][ybstride], on the box
 *          box[0..2] + [0, box[3]) x [0, box[4]) x [0, box[5]) (x, y, z),
 *          which must hold those cells; 0 when yb is NULL)
 *        + s[i][c]                              (the far-field operator's
 *          value, float32 [..][sstride]; 0 when s is NULL),
 *   where mask[i] != 0 (uint8 per cell, may be NULL): v = 0 (mask_all) or
 *   the s term left out (!mask_all);
 *   fields[c][i] += gain * scale[c] * v[c].                               */
int conv3d_apply_closure(int32_t nx, int32_t ny, int32_t nz, float* const* fields, const float* s,
                         int32_t sstride, const float* yb, const float* y0b, int32_t ybstride,
                         const int32_t* box, const unsigned char* mask, int32_t mask_all,
                         float net_gain, const float* scale, float gain, int32_t margin,
                         void* stream);
/* The near-source patch's correction of one interval: fields[c] += gain *
 * corr[c] (fused multiply-add) over the box (x0, y0, z0) + [0, bnx) x
 * [0, bny) x [0, bnz); corr float32 [6][bnz][bny][bnx].                  */
int conv3d_patch_add(int32_t nx, int32_t ny, int32_t nz, float* const* fields, const float* corr,
                     int32_t x0, int32_t y0, int32_t z0, int32_t bnx, int32_t bny, int32_t bnz,
                     float gain, void* stream);

/* The closure's near-source patch (no reference counterpart): a local grid
 * and its 2x-refined twin run side by side, and the coarse box of side^3
 * cells at coarse (x0, y0, z0) is kept equal to the restricted fine state.
 * Fine fields are float32, x fastest, nxf x nyf x nzf with even sides; fine
 * index of coarse (K, J, I) is (2K, 2J, 2I).  Restriction: coarse Ex is the
 * mean of fine Ex at x and x+1 (Ey, Ez likewise along their own axes);
 * coarse Hx the mean of the four fine Hx at (y, z) + {0,1}^2 (Hy, Hz
 * likewise).
 * conv3d_patch_restrict_h: out[3][side^3] = the restricted fine H (hf: 3
 * device pointers), kept for the mean with the next fine H.
 * conv3d_patch_correct: st = restricted fine E (ef) and the mean of hprev and
 * the restricted fine H (hf); for each of the six coarse fields (coarse: 6
 * device pointers, nxc x nyc x nzc), corr[c][i] = st - coarse and coarse +=
 * corr over the box.                                                       */
int conv3d_patch_restrict_h(const float* const* hf, int32_t nxf, int32_t nyf, int32_t nzf,
                            int32_t x0, int32_t y0, int32_t z0, int32_t side, float* out,
                            void* stream);
int conv3d_patch_correct(const float* const* ef, const float* const* hf, const float* hprev,
                         int32_t nxf, int32_t nyf, int32_t nzf, float* const* coarse,
                         int32_t nxc, int32_t nyc, int32_t nzc, int32_t x0, int32_t y0,
                         int32_t z0, int32_t side, float* corr, void* stream);

const char* conv3d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* JAXPHYS_AMD_CONV3D_H */
