// Learned sshex8 smoother: a per-macro-element 3x3x3 CNN on the (L+1)^3 micro-grid (zero
// padding outside the element).  Training kernels run in FP32 (direct convolution with the
// element's channels staged in LDS); inference runs the fused network on MFMA with BF16
// weights and activations and FP32 accumulation.
//
// Layouts: activations [E][C][nxe] float (node index = lidx), weights [Cout][Cin][27] with
// offset k = (dz+1)*9 + (dy+1)*3 + (dx+1).
#include "cnn_kernels.hpp"
#include "common.hpp"

namespace sfem_amd {

__device__ __forceinline__ int nbr(int L, int x, int y, int z, int k) {
    const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
    const int xx = x + dx, yy = y + dy, zz = z + dz;
    if (xx < 0 || xx > L || yy < 0 || yy > L || zz < 0 || zz > L) return -1;
    return lidx(L, xx, yy, zz);
}

// out[e][co][v] = sum_ci sum_k W[co][ci][k] act(in[e][ci][v + off_k]);  act = relu if relu_in
__global__ void __launch_bounds__(256) conv_fwd_kernel(const int L, const ptrdiff_t E, const int Cin, const int Cout,
                                                       const float *__restrict__ in, const float *__restrict__ W,
                                                       const int relu_in, float *__restrict__ out) {
    extern __shared__ float smem[];
    const int nxe = nxe_of(L);
    float *sin = smem;  // [Cin][nxe]; the weights (<= 28 KB) are read through the caches
    const float *sw = W;
    const ptrdiff_t e = blockIdx.x;
    if (e >= E) return;  // uniform per block
    for (int i = threadIdx.x; i < Cin * nxe; i += blockDim.x) {
        float v = in[(size_t)e * Cin * nxe + i];
        sin[i] = relu_in ? fmaxf(v, 0.f) : v;
    }
    __syncthreads();
    const int N1 = L + 1;
    // blockIdx.y splits the element's outputs over gridDim.y blocks (each stages the whole input)
    for (int t = blockIdx.y * blockDim.x + threadIdx.x; t < Cout * nxe; t += blockDim.x * gridDim.y) {
        const int co = t / nxe, v = t - co * nxe;
        const int x = v % N1, y = (v / N1) % N1, z = v / (N1 * N1);
        int nb[27];
#pragma unroll
        for (int k = 0; k < 27; k++) nb[k] = nbr(L, x, y, z, k);
        float acc = 0.f;
        for (int ci = 0; ci < Cin; ci++) {
            const float *w = sw + ((size_t)co * Cin + ci) * 27;
            const float *a = sin + (size_t)ci * nxe;
#pragma unroll
            for (int k = 0; k < 27; k++)
                if (nb[k] >= 0) acc += w[k] * a[nb[k]];
        }
        out[((size_t)e * Cout + co) * nxe + v] = acc;
    }
}

// din[e][ci][v] = relu'(pre) * sum_co sum_k W[co][ci][k] dout[e][co][v - off_k]
// (act_in holds the layer input before the activation when relu_in, for the mask)
__global__ void __launch_bounds__(256) conv_bwd_data_kernel(const int L, const ptrdiff_t E, const int Cin,
                                                            const int Cout, const float *__restrict__ dout,
                                                            const float *__restrict__ W,
                                                            const float *__restrict__ act_in, const int relu_in,
                                                            float *__restrict__ din) {
    extern __shared__ float smem[];
    const int nxe = nxe_of(L);
    float *sd = smem;  // [Cout][nxe]
    const float *sw = W;
    const ptrdiff_t e = blockIdx.x;
    if (e >= E) return;
    for (int i = threadIdx.x; i < Cout * nxe; i += blockDim.x) sd[i] = dout[(size_t)e * Cout * nxe + i];
    __syncthreads();
    const int N1 = L + 1;
    for (int t = blockIdx.y * blockDim.x + threadIdx.x; t < Cin * nxe; t += blockDim.x * gridDim.y) {
        const int ci = t / nxe, v = t - ci * nxe;
        const int x = v % N1, y = (v / N1) % N1, z = v / (N1 * N1);
        float acc = 0.f;
        // out[m] += W[k] in[m + off_k]  =>  din[v] += W[k] dout[v - off_k]
#pragma unroll
        for (int k = 0; k < 27; k++) {
            const int m = nbr(L, x, y, z, 26 - k);  // v - off_k
            if (m < 0) continue;
            for (int co = 0; co < Cout; co++) acc += sw[((size_t)co * Cin + ci) * 27 + k] * sd[(size_t)co * nxe + m];
        }
        if (relu_in && act_in[((size_t)e * Cin + ci) * nxe + v] <= 0.f) acc = 0.f;
        din[((size_t)e * Cin + ci) * nxe + v] = acc;
    }
}

// partial[b][co][ci][k] = sum over the block's elements and nodes of dout[co][v] act(in[ci][v+off_k]);
// blockIdx.y selects a chunk of kCoChunk output channels (LDS: Cin + kCoChunk channels).
constexpr int kCoChunk = 4;
__global__ void __launch_bounds__(256) conv_bwd_weight_kernel(const int L, const ptrdiff_t E, const int Cin,
                                                              const int Cout, const float *__restrict__ dout,
                                                              const float *__restrict__ in, const int relu_in,
                                                              float *__restrict__ partial) {
    extern __shared__ float smem[];
    const int nxe = nxe_of(L);
    const int co0 = blockIdx.y * kCoChunk;
    const int nco = min(kCoChunk, Cout - co0);
    float *sin = smem;                     // [Cin][nxe]
    float *sd = smem + (size_t)Cin * nxe;   // [nco][nxe]
    const int nw_chunk = nco * Cin * 27;     // <= 4 * 16 * 27 = 1728 -> <= 7 per thread
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = 0.f;
    for (ptrdiff_t e = blockIdx.x; e < E; e += gridDim.x) {
        __syncthreads();
        for (int i = threadIdx.x; i < Cin * nxe; i += blockDim.x) {
            const float v = in[(size_t)e * Cin * nxe + i];
            sin[i] = relu_in ? fmaxf(v, 0.f) : v;
        }
        for (int i = threadIdx.x; i < nco * nxe; i += blockDim.x)
            sd[i] = dout[((size_t)e * Cout + co0) * nxe + i];
        __syncthreads();
        for (int j = 0; j < 8; j++) {
            const int t = threadIdx.x + j * blockDim.x;
            if (t >= nw_chunk) break;
            const int col = t / (Cin * 27), r = t - col * Cin * 27, ci = r / 27, k = r - ci * 27;
            const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
            const float *d = sd + (size_t)col * nxe;
            const float *a = sin + (size_t)ci * nxe;
            float s = 0.f;
            for (int z = (dz < 0 ? 1 : 0); z <= L - (dz > 0 ? 1 : 0); z++)
                for (int y = (dy < 0 ? 1 : 0); y <= L - (dy > 0 ? 1 : 0); y++)
                    for (int x = (dx < 0 ? 1 : 0); x <= L - (dx > 0 ? 1 : 0); x++)
                        s += d[lidx(L, x, y, z)] * a[lidx(L, x + dx, y + dy, z + dz)];
            acc[j] += s;
        }
    }
    const int nw = Cout * Cin * 27;
    for (int j = 0; j < 8; j++) {
        const int t = threadIdx.x + j * blockDim.x;
        if (t >= nw_chunk) break;
        partial[(size_t)blockIdx.x * nw + (size_t)co0 * Cin * 27 + t] = acc[j];
    }
}

__global__ void sum_partials_f32_kernel(const int nb, const int nw, const float *__restrict__ partial,
                                        const float scale, float *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nw) return;
    float s = 0.f;
    for (int b = 0; b < nb; b++) s += partial[(size_t)b * nw + t];
    out[t] += scale * s;
}

// out = a - b (the odd symmetrisation g(z) - g(-z))
__global__ void sub_halves_kernel(const ptrdiff_t n, const float *__restrict__ a, const float *__restrict__ b,
                                  float *__restrict__ out) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        out[i] = a[i] - b[i];
}

// x2 = [x; -x]
__global__ void make_pm_kernel(const ptrdiff_t n, const float *__restrict__ x, float *__restrict__ x2) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        x2[i] = x[i];
        x2[n + i] = -x[i];
    }
}

// squared-error partials: part[b] = (sum (o - t)^2, sum t^2) over a fixed grid
__global__ void __launch_bounds__(256) sq_err_partial_kernel(const ptrdiff_t n, const float *__restrict__ o,
                                                             const float *__restrict__ t, double *__restrict__ part) {
    __shared__ double s1[256], s2[256];
    double a = 0, b = 0;
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        const double d = (double)o[i] - (double)t[i];
        a += d * d;
        b += (double)t[i] * (double)t[i];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            s1[threadIdx.x] += s1[threadIdx.x + st];
            s2[threadIdx.x] += s2[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

// dpm = [2 (o - t) / norm; -2 (o - t) / norm]
__global__ void dloss_kernel(const ptrdiff_t n, const float *__restrict__ o, const float *__restrict__ t,
                             const float inv_norm, float *__restrict__ dpm) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        const float d = 2.f * (o[i] - t[i]) * inv_norm;
        dpm[i] = d;
        dpm[n + i] = -d;
    }
}

__global__ void adam_kernel(const int n, const float *__restrict__ g, float *__restrict__ w, float *__restrict__ m,
                            float *__restrict__ v, const float lr, const float b1, const float b2, const float eps,
                            const float c1, const float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// out[e * ld + off + v] = vec[idx[e][v]*bs + comp] * (scale ? scale[...] : 1)  (ld = nxe, off = 0: [E][nxe];
// ld = bs * nxe, off = comp * nxe: the channel comp of [E][bs][nxe])
template <typename T>
__global__ void gather_elements_kernel(const ptrdiff_t E, const int nxe, const int32_t *__restrict__ idx, const int bs,
                                       const int comp, const T *__restrict__ vec, const T *__restrict__ scale,
                                       const int ld, const int off, float *__restrict__ out) {
    const ptrdiff_t n = E * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t d = (ptrdiff_t)idx[t] * bs + comp;
        const ptrdiff_t e = t / nxe;
        out[e * ld + off + (t - e * nxe)] = (float)(scale ? vec[d] * scale[d] : vec[d]);
    }
}

// vec[idx[e][v]*bs + comp] += alpha * val[e * ld + off + v] / count[idx]
template <typename T>
__global__ void scatter_elements_kernel(const ptrdiff_t E, const int nxe, const int32_t *__restrict__ idx, const int bs,
                                        const int comp, const uint16_t *__restrict__ count, const float alpha,
                                        const int ld, const int off, const float *__restrict__ val,
                                        T *__restrict__ vec) {
    const ptrdiff_t n = E * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const int32_t g = idx[t];
        const ptrdiff_t e = t / nxe;
        atomicAdd(&vec[(ptrdiff_t)g * bs + comp], (T)(alpha * val[e * ld + off + (t - e * nxe)] / (float)count[g]));
    }
}

// ------------------------------------------------------------------ launchers
static size_t smem_fwd(int L, int Cin, int /*Cout*/) { return sizeof(float) * ((size_t)Cin * nxe_of(L)); }

// blocks per element for n outputs of 256-thread blocks (training batches hold few elements)
static unsigned conv_split(int n) {
    const int b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 16 ? 16 : b));
}

int cnn_conv_forward(int L, ptrdiff_t E, int Cin, int Cout, const float *in, const float *W, int relu_in, float *out,
                     hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    conv_fwd_kernel<<<dim3((unsigned)E, conv_split(Cout * nxe_of(L))), 256, smem_fwd(L, Cin, Cout), s>>>(
        L, E, Cin, Cout, in, W, relu_in, out);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int cnn_conv_backward_data(int L, ptrdiff_t E, int Cin, int Cout, const float *dout, const float *W, const float *act_in,
                           int relu_in, float *din, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    const size_t smem = sizeof(float) * ((size_t)Cout * nxe_of(L));
    conv_bwd_data_kernel<<<dim3((unsigned)E, conv_split(Cin * nxe_of(L))), 256, smem, s>>>(L, E, Cin, Cout, dout, W,
                                                                                         act_in, relu_in, din);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int cnn_conv_backward_weights(int L, ptrdiff_t E, int Cin, int Cout, const float *dout, const float *in, int relu_in,
                              float *partial, int nblocks, float *dW, hipStream_t s) {
    const int nw = Cout * Cin * 27;
    if (kCoChunk * Cin * 27 > 8 * 256) {
        set_last_error("cnn: layer too wide for the weight-gradient kernel");
        return SFEM_FAILURE;
    }
    if (E == 0) return SFEM_SUCCESS;
    const size_t smem = sizeof(float) * ((size_t)(Cin + kCoChunk) * nxe_of(L));
    dim3 grid(nblocks, (Cout + kCoChunk - 1) / kCoChunk);
    conv_bwd_weight_kernel<<<grid, 256, smem, s>>>(L, E, Cin, Cout, dout, in, relu_in, partial);
    sum_partials_f32_kernel<<<(nw + 255) / 256, 256, 0, s>>>(nblocks, nw, partial, 1.f, dW);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int cnn_make_pm(ptrdiff_t n, const float *x, float *x2, hipStream_t s) {
    make_pm_kernel<<<grid_for(n), kBlock, 0, s>>>(n, x, x2);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int cnn_sub_halves(ptrdiff_t n, const float *a, const float *b, float *out, hipStream_t s) {
    sub_halves_kernel<<<grid_for(n), kBlock, 0, s>>>(n, a, b, out);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int cnn_sq_err(ptrdiff_t n, const float *o, const float *t, double *ws, double *err_host, double *norm_host,
               hipStream_t s) {
    const int nb = 256;
    sq_err_partial_kernel<<<nb, 256, 0, s>>>(n, o, t, ws);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    double h[2 * 256];
    SFEM_AMD_HIP_CHECK(hipMemcpyAsync(h, ws, sizeof(h), hipMemcpyDeviceToHost, s));
    SFEM_AMD_HIP_CHECK(hipStreamSynchronize(s));
    double a = 0, b = 0;
    for (int i = 0; i < nb; i++) {
        a += h[2 * i];
        b += h[2 * i + 1];
    }
    *err_host = a;
    *norm_host = b;
    return SFEM_SUCCESS;
}

int cnn_dloss(ptrdiff_t n, const float *o, const float *t, float inv_norm, float *dpm, hipStream_t s) {
    dloss_kernel<<<grid_for(n), kBlock, 0, s>>>(n, o, t, inv_norm, dpm);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int cnn_adam(int n, const float *g, float *w, float *m, float *v, float lr, float b1, float b2, float eps, int step,
             hipStream_t s) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    adam_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, g, w, m, v, lr, b1, b2, eps, c1, c2);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
int gather_elements(ptrdiff_t E, int nxe, const int32_t *idx, int bs, int comp, const T *vec, const T *scale,
                    float *out, hipStream_t s, bool channels) {
    if (E == 0) return SFEM_SUCCESS;
    gather_elements_kernel<T><<<grid_for(E * nxe), kBlock, 0, s>>>(E, nxe, idx, bs, comp, vec, scale,
                                                                   channels ? bs * nxe : nxe, channels ? comp * nxe : 0,
                                                                   out);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
int scatter_elements(ptrdiff_t E, int nxe, const int32_t *idx, int bs, int comp, const uint16_t *count, float alpha,
                     const float *val, T *vec, hipStream_t s, bool channels) {
    if (E == 0) return SFEM_SUCCESS;
    scatter_elements_kernel<T><<<grid_for(E * nxe), kBlock, 0, s>>>(E, nxe, idx, bs, comp, count, alpha,
                                                                    channels ? bs * nxe : nxe,
                                                                    channels ? comp * nxe : 0, val, vec);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template int gather_elements<double>(ptrdiff_t, int, const int32_t *, int, int, const double *, const double *, float *,
                                     hipStream_t, bool);
template int gather_elements<float>(ptrdiff_t, int, const int32_t *, int, int, const float *, const float *, float *,
                                    hipStream_t, bool);
template int scatter_elements<double>(ptrdiff_t, int, const int32_t *, int, int, const uint16_t *, float, const float *,
                                      double *, hipStream_t, bool);
template int scatter_elements<float>(ptrdiff_t, int, const int32_t *, int, int, const uint16_t *, float, const float *,
                                     float *, hipStream_t, bool);

// ------------------------------------------------------------------ fused BF16 MFMA inference
// One macro element per block (512 threads = 8 waves); the element's
// activations stay in LDS for all four layers and both signs of f(z) = g(z) - g(-z).  Nodes are
// processed in tiles of 16 (the MFMA N dimension); per tile and layer:
//   layer 1 (1 -> 16):  D[co][n]  = sum_{k<32} W1[co][k] z[n + off_k]         one 16x16x32 MFMA
//   layer 2,3 (16->16): D[co][n] += sum_{pairs p} W[co][ci,off] act[n + off]  14 16x16x32 MFMAs,
//                       K = 32 = two offsets x 16 input channels; each wave runs three tiles at
//                       once (three independent accumulation chains)
//   layer 4 (16 -> 1):  P[k][n]   = sum_ci W4[ci][k] act[ci][n] (16x16x16 MFMAs, 16 offsets k per
//                       pass, two passes), then out[n] = sum_k P[k][n + off_k] from LDS.
// Lane layouts (64 lanes, g = lane / 16, c = lane % 16):
//   16x16x32 A: row c, k = 8g + j (j < 8);  B: column c, k = 8g + j;  D: column c, rows 4g + r
//   16x16x16 A: row c, k = 4g + j (j < 4);  B: column c, k = 4g + j;  D: as above
// The input and the activations of layers 1-3 live on the element's grid with a one-point zero
// halo ((L+3)^3 points), so neighbour reads need no validity tests: one address per tile and a
// per-pair offset.  Activations are stored post-ReLU in BF16 as two channel planes
// [half][point][8 channels] (16-byte B reads).  Layer 4's stencil sum reads the unpadded P, with
// the padding node NXE (P == 0) standing in for missing neighbours.
typedef short mf_s4 __attribute__((ext_vector_type(4)));
typedef __bf16 mf_b8 __attribute__((ext_vector_type(8)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int mf_u4 __attribute__((ext_vector_type(4)));
typedef unsigned int mf_u2 __attribute__((ext_vector_type(2)));

typedef __bf16 mf_b2 __attribute__((ext_vector_type(2)));
typedef float mf_f2 __attribute__((ext_vector_type(2)));

// float -> bf16, round to nearest even: the conversions lower to gfx950's v_cvt_pk_bf16_f32
// (bf_pack converts two values with one instruction)
__device__ __forceinline__ __bf16 f2bf(float f) { return (__bf16)f; }
__device__ __forceinline__ unsigned bf_pack(float lo, float hi) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(mf_f2{lo, hi}, mf_b2));
}

template <int L>
struct CnnShape {
    static constexpr int N1 = L + 1;
    static constexpr int NXE = N1 * N1 * N1;
    // node tiles of 16 with at least one padding node (node NXE: P[k][NXE] == 0 stands in for the
    // missing neighbours of the layer-4 stencil sum)
    static constexpr int T = NXE / 16 + 1;
    static constexpr int NP = T * 16;
    // layers 1-3 hold the element on grids with a zero halo, so every node's 27 neighbours are in
    // range and read without tests: the input on the (L+3)^3 grid, the activations on a grid with
    // line stride SY and plane stride SZ chosen so that consecutive nodes, also across line and
    // plane ends, sit one 16-byte row apart mod 8 (eight lanes read all 32 LDS banks)
    static constexpr int P1 = L + 3;
    static constexpr int NH = P1 * P1 * P1;
    static constexpr int SY = L + 9;  // >= L + 3, = L + 1 (mod 8)
    static constexpr int SZ = P1 * SY + ((((L * SY + L + 1) - P1 * SY) % 8) + 8) % 8;  // = L SY + L + 1 (mod 8)
    static constexpr int NA = P1 * SZ;
    static constexpr int THREADS = 512;
    static constexpr int TPG = 3;  // tiles per wave iteration in the hidden layers
    // zs [NH] float | actX [2][NA][8] bf16 | actY [2][NA][8] bf16, aliased by P [16][NP] float
    // (actY is cleared again at the start of each sign)
    static constexpr size_t ZS = (sizeof(float) * NH + 15) / 16 * 16;
    static constexpr size_t ACT = 32 * (size_t)NA;
    static constexpr size_t LDS = ZS + ACT + (ACT > 64 * (size_t)NP ? ACT : 64 * (size_t)NP);
};

// Neighbour validity as bit tests.  cnn_avail: the directions in which node (x, y, z) has a
// neighbour (bits x-, x+, y-, y+, z-, z+) and bit 6 for a real node (in); cnn_need(o): the bits offset
// o requires (bit 7, never available, for the padding offsets o >= 27).  Offset o of the node exists
// iff (cnn_need(o) & ~cnn_avail) == 0.
template <int L>
__device__ __forceinline__ unsigned cnn_avail(int x, int y, int z, bool in) {
    return (unsigned)(x > 0) | (unsigned)(x < L) << 1 | (unsigned)(y > 0) << 2 | (unsigned)(y < L) << 3 |
           (unsigned)(z > 0) << 4 | (unsigned)(z < L) << 5 | (unsigned)in << 6;
}
__host__ __device__ constexpr unsigned cnn_need(int o) {
    return o >= 27 ? 128u
                   : 64u | (o % 3 == 0 ? 1u : 0u) | (o % 3 == 2 ? 2u : 0u) | ((o / 3) % 3 == 0 ? 4u : 0u) |
                         ((o / 3) % 3 == 2 ? 8u : 0u) | (o / 9 == 0 ? 16u : 0u) | (o / 9 == 2 ? 32u : 0u);
}

template <int L>
__device__ __forceinline__ int cnn_shift(int o) {
    constexpr int N1 = L + 1;
    return (o / 9 - 1) * N1 * N1 + ((o / 3) % 3 - 1) * N1 + (o % 3 - 1);
}

// node n = lidx(x, y, z) on a haloed grid with line stride SY and plane stride SZ;
// cnn_pshift(o): offset o there (0 for the zero-weight padding offsets o >= 27, which then read
// the node itself)
template <int L, int SY, int SZ>
__device__ __forceinline__ int cnn_pad(int n) {
    constexpr int N1 = L + 1;
    const int x = n % N1, y = (n / N1) % N1, z = n / (N1 * N1);
    return (x + 1) + (y + 1) * SY + (z + 1) * SZ;
}
template <int SY, int SZ>
__device__ __forceinline__ int cnn_pshift(int o) {
    return o >= 27 ? 0 : (o / 9 - 1) * SZ + ((o / 3) % 3 - 1) * SY + (o % 3 - 1);
}

// 4 channels (4g .. 4g+3) of grid point q, as stored from an MFMA D fragment
__device__ __forceinline__ unsigned short *act_quad(unsigned short *a, int NA, int q, int g) {
    return a + (g >> 1) * NA * 8 + q * 8 + 4 * (g & 1);
}

__device__ __forceinline__ mf_u2 relu_pack(const mf_f4 &acc, float sg) {
    mf_u2 o2;
    o2[0] = bf_pack(fmaxf(sg * acc[0], 0.f), fmaxf(sg * acc[1], 0.f));
    o2[1] = bf_pack(fmaxf(sg * acc[2], 0.f), fmaxf(sg * acc[3], 0.f));
    return o2;
}

template <int L>
__global__ void __launch_bounds__(512, 2) cnn_infer_kernel(const ptrdiff_t E, const unsigned short *__restrict__ W16,
                                                           const float *__restrict__ zin, float *__restrict__ out) {
    using S = CnnShape<L>;
    constexpr int NP = S::NP, NH = S::NH, NA = S::NA, P1 = S::P1, SY = S::SY, SZ = S::SZ, TPG = S::TPG,
                  NB = (S::NXE + S::THREADS - 1) / S::THREADS;
    extern __shared__ __align__(16) unsigned char cnn_smem[];
    float *zs = (float *)cnn_smem;
    unsigned short *actX = (unsigned short *)(cnn_smem + S::ZS);
    unsigned short *actY = actX + 16 * NA;
    float *P = (float *)actY;
    const ptrdiff_t e = blockIdx.x;
    if (e >= E) return;  // uniform per block
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = S::THREADS / 64;
    const int g = lane >> 4, c = lane & 15;

    // the input on its haloed grid; actX zeroed (the layers write only the interior points, so its
    // halo stays zero for both signs)
    {
        // all of this thread's input loads at once (halo and tail points read the element's first
        // value, a valid address, and store zero)
        constexpr int NBI = (NH + S::THREADS - 1) / S::THREADS;
        float zv[NBI];
#pragma unroll
        for (int b = 0; b < NBI; b++) {
            const int q = tid + b * S::THREADS;
            const int px = q % P1 - 1, py = (q / P1) % P1 - 1, pz = q / (P1 * P1) - 1;
            const bool in = q < NH && (unsigned)px <= (unsigned)L && (unsigned)py <= (unsigned)L &&
                            (unsigned)pz <= (unsigned)L;
            const float v = zin[(size_t)e * S::NXE + (in ? (pz * S::N1 + py) * S::N1 + px : 0)];
            zv[b] = in ? v : 0.f;
        }
#pragma unroll
        for (int b = 0; b < NBI; b++) {
            const int q = tid + b * S::THREADS;
            if (q < NH) zs[q] = zv[b];
        }
    }
    for (int i = tid; i < 2 * NA; i += S::THREADS) ((mf_u4 *)actX)[i] = mf_u4{0u, 0u, 0u, 0u};
    // weights in lane order: layer 1 [64][8], layers 2-3 [14][64][8], layer 4 [2][64][4]
    const mf_u4 w1 = *(const mf_u4 *)(W16 + lane * 8);
    int sh1[8];  // layer 1: lane group g gathers offsets 8g .. 8g+7
#pragma unroll
    for (int j = 0; j < 8; j++) sh1[j] = cnn_pshift<P1, P1 * P1>(8 * g + j);
    const bool hi = (g >> 1) != 0;  // hidden layers: lane groups 2, 3 take the pair's second offset
    float gplus[NB];
    __syncthreads();

    for (int sign = 0; sign < 2; sign++) {
        const float sg = sign ? -1.f : 1.f;
        // ---- layer 1: zs -> actX; actY (P of the previous sign) cleared for layer 2
        for (int i = tid; i < 2 * NA; i += S::THREADS) ((mf_u4 *)actY)[i] = mf_u4{0u, 0u, 0u, 0u};
        for (int t = wave; t < S::T; t += nw) {
            const int n = t * 16 + c;
            const bool real = n < S::NXE;
            const int qz = cnn_pad<L, P1, P1 * P1>(real ? n : 0);
            mf_b8 b;
#pragma unroll
            for (int j = 0; j < 8; j++) b[j] = f2bf(zs[qz + sh1[j]]);
            mf_f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mf_b8, w1), b, acc, 0, 0, 0);
            if (real) *(mf_u2 *)act_quad(actX, NA, cnn_pad<L, SY, SZ>(n), g) = relu_pack(acc, sg);
        }
        __syncthreads();
        // ---- layers 2 and 3: actX -> actY -> actX
#pragma unroll 1
        for (int layer = 0; layer < 2; layer++) {
            const unsigned short *ain = layer ? actY : actX;
            unsigned short *aout = layer ? actX : actY;
            const unsigned short *wl = W16 + 512 + layer * 7168;
            mf_u4 w[14];
#pragma unroll
            for (int p = 0; p < 14; p++) w[p] = *(const mf_u4 *)(wl + (p * 64 + lane) * 8);
            // byte addresses: the lane's channel plane plus 16 bytes per grid point
            const unsigned char *plane = (const unsigned char *)(ain + (g & 1) * NA * 8);
            for (int t = TPG * wave; t < S::T; t += TPG * nw) {
                int q[TPG], qb[TPG];
                bool real[TPG];
                mf_f4 acc[TPG];
#pragma unroll
                for (int h = 0; h < TPG; h++) {
                    const int n = (t + h) * 16 + c;
                    real[h] = n < S::NXE;
                    q[h] = cnn_pad<L, SY, SZ>(real[h] ? n : 0);
                    qb[h] = q[h] * 16;
                    acc[h] = mf_f4{0.f, 0.f, 0.f, 0.f};
                }
#pragma unroll
                for (int p = 0; p < 14; p++) {
                    const int shb = 16 * (hi ? cnn_pshift<SY, SZ>(2 * p + 1) : cnn_pshift<SY, SZ>(2 * p));
                    mf_u4 bv[TPG];
#pragma unroll
                    for (int h = 0; h < TPG; h++) bv[h] = *(const mf_u4 *)(plane + qb[h] + shb);
#pragma unroll
                    for (int h = 0; h < TPG; h++)
                        acc[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mf_b8, w[p]),
                                                                          __builtin_bit_cast(mf_b8, bv[h]), acc[h], 0,
                                                                          0, 0);
                }
#pragma unroll
                for (int h = 0; h < TPG; h++)
                    if (real[h]) *(mf_u2 *)act_quad(aout, NA, q[h], g) = relu_pack(acc[h], 1.f);
            }
            __syncthreads();
        }
        // ---- layer 4 in two passes of 16 offsets: P[r][n] = sum_ci W4[ci][16 pass + r] act[ci][n],
        // then v[n] += sum_r P[r][n + off_{16 pass + r}]
        float v[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) v[i] = 0.f;
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            const mf_u2 w4 = *(const mf_u2 *)(W16 + 512 + 2 * 7168 + pass * 256 + lane * 4);
            for (int t = wave; t < S::T; t += nw) {
                const int n = t * 16 + c;
                const bool real = n < S::NXE;
                const mf_u2 ld = *(const mf_u2 *)act_quad(actX, NA, cnn_pad<L, SY, SZ>(real ? n : 0), g);
                const mf_u2 bv = real ? ld : mf_u2{0u, 0u};  // padding nodes: P[r][n] = 0
                mf_f4 a = {0.f, 0.f, 0.f, 0.f};
                a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(mf_s4, w4), __builtin_bit_cast(mf_s4, bv),
                                                               a, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) P[(4 * g + r) * NP + n] = a[r];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < NB; i++) {
                const int n = tid + i * S::THREADS;
                if (n < S::NXE) {
                    const unsigned na = ~cnn_avail<L>(n % S::N1, (n / S::N1) % S::N1, n / (S::N1 * S::N1), true);
                    // missing neighbours add P[r][NXE] == 0 (the padding node's product)
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int k = 16 * pass + r;
                        if (k < 27) v[i] += P[r * NP + ((cnn_need(k) & na) == 0 ? n + cnn_shift<L>(k) : S::NXE)];
                    }
                }
            }
            __syncthreads();  // P is rewritten by the next pass / sign
        }
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int n = tid + i * S::THREADS;
            if (n < S::NXE) {
                if (sign == 0)
                    gplus[i] = v[i];
                else
                    out[(size_t)e * S::NXE + n] = gplus[i] - v[i];
            }
        }
    }
}

template <int L>
static int launch_cnn_infer(ptrdiff_t E, const uint16_t *W16, const float *z, float *out, hipStream_t s) {
    using S = CnnShape<L>;
    static bool attr = false;
    if (!attr) {
        SFEM_AMD_HIP_CHECK(
            hipFuncSetAttribute((const void *)cnn_infer_kernel<L>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::LDS));
        attr = true;
    }
    cnn_infer_kernel<L><<<(unsigned)E, S::THREADS, S::LDS, s>>>(E, (const unsigned short *)W16, z, out);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int cnn_infer_mfma(int L, ptrdiff_t E, const uint16_t *W16, const float *z, float *out, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    switch (L) {
        case 1: return launch_cnn_infer<1>(E, W16, z, out, s);
        case 2: return launch_cnn_infer<2>(E, W16, z, out, s);
        case 3: return launch_cnn_infer<3>(E, W16, z, out, s);
        case 4: return launch_cnn_infer<4>(E, W16, z, out, s);
        case 5: return launch_cnn_infer<5>(E, W16, z, out, s);
        case 6: return launch_cnn_infer<6>(E, W16, z, out, s);
        case 7: return launch_cnn_infer<7>(E, W16, z, out, s);
        case 8: return launch_cnn_infer<8>(E, W16, z, out, s);
        default: set_last_error("cnn_infer_mfma: level must be 1..8"); return SFEM_FAILURE;
    }
}

}  // namespace sfem_amd
