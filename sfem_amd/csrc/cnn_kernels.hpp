// Host-callable launchers of cnn_kernels.hip (learned smoother).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sfem_amd {
int cnn_conv_forward(int L, ptrdiff_t E, int Cin, int Cout, const float *in, const float *W, int relu_in, float *out,
                     hipStream_t s);
int cnn_conv_backward_data(int L, ptrdiff_t E, int Cin, int Cout, const float *dout, const float *W, const float *act_in,
                           int relu_in, float *din, hipStream_t s);
// dW += sum of per-block partials (partial: nblocks * Cout*Cin*27 floats)
int cnn_conv_backward_weights(int L, ptrdiff_t E, int Cin, int Cout, const float *dout, const float *in, int relu_in,
                              float *partial, int nblocks, float *dW, hipStream_t s);
int cnn_make_pm(ptrdiff_t n, const float *x, float *x2, hipStream_t s);
int cnn_sub_halves(ptrdiff_t n, const float *a, const float *b, float *out, hipStream_t s);
// ws: 512 doubles of device workspace; returns sum (o-t)^2 and sum t^2 (synchronising)
int cnn_sq_err(ptrdiff_t n, const float *o, const float *t, double *ws, double *err_host, double *norm_host,
               hipStream_t s);
int cnn_dloss(ptrdiff_t n, const float *o, const float *t, float inv_norm, float *dpm, hipStream_t s);
int cnn_adam(int n, const float *g, float *w, float *m, float *v, float lr, float b1, float b2, float eps, int step,
             hipStream_t s);
// Element values of component comp of a bs-interleaved vector: out [E][nxe], or with channels the
// channel comp of out [E][bs][nxe]
template <typename T>
int gather_elements(ptrdiff_t E, int nxe, const int32_t *idx, int bs, int comp, const T *vec, const T *scale,
                    float *out, hipStream_t s, bool channels = false);
// vec[node * bs + comp] += alpha * val / count[node] over the element values (layout as gather_elements)
template <typename T>
int scatter_elements(ptrdiff_t E, int nxe, const int32_t *idx, int bs, int comp, const uint16_t *count, float alpha,
                     const float *val, T *vec, hipStream_t s, bool channels = false);
// fused BF16 MFMA inference (hidden 16, 4 layers, L <= 8); W16: bf16 weights of the 4 layers in
// MFMA lane order (kCnnPackedWeights entries, see cnn_pack_weights)
constexpr int kCnnPackedWeights = 512 + 2 * 7168 + 512;
void cnn_pack_weights(const float *W, uint16_t *packed);
int cnn_infer_mfma(int L, ptrdiff_t E, const uint16_t *W16, const float *z, float *out, hipStream_t s);
}  // namespace sfem_amd
