// Learned smoother (chosen form 4): the CNN object (weights, Adam state, training step) and
// the smoother operator x += damping * f(D^-1 (b - A x)) assembled per macro element.
#include <cmath>
#include <random>
#include <vector>

#include "cnn_kernels.hpp"
#include "internal.hpp"

using namespace sfem_amd;

struct sfem_amd_cnn {
    int hidden{16}, layers{4};
    int io{1};                    // input and output channels: 1, or 3 (one per displacement component)
    std::vector<int> chans;       // layer l: chans[l] -> chans[l+1]
    std::vector<size_t> offset;   // weight offset of layer l
    size_t nparams{0};
    DevArray<float> W, grad, m, v;
    DevArray<uint16_t> W16;       // bf16 copy for MFMA inference
    bool w16_valid{false};
    int step{0};
    // workspaces (grown on demand)
    DevArray<float> x2, dpm, out2, part;
    std::vector<DevArray<float> *> acts;  // pre-activations of each layer output, [2E][C][nxe]
    DevArray<float> dA, dB;
    DevArray<double> ws;
    ~sfem_amd_cnn() {
        for (auto *a : acts) delete a;
    }
};

namespace {

int ensure(DevArray<float> &a, size_t n) {
    if (a.n >= n) return SFEM_SUCCESS;
    return a.alloc(n);
}

int check_level(int L) {
    if (L < 1 || L > 8) {
        set_last_error("learned smoother: micro-grid level must be 1..8 (LDS-staged element tiles)");
        return SFEM_FAILURE;
    }
    return SFEM_SUCCESS;
}

// FP32 forward of g on the batch x2 ([2E][io][nxe]); keeps every layer's output in acts.
int forward_g(sfem_amd_cnn *c, int L, ptrdiff_t E2, hipStream_t s) {
    const size_t nxe = nxe_of(L);
    const float *in = c->x2.p;
    for (int l = 0; l < c->layers; l++) {
        const int ci = c->chans[l], co = c->chans[l + 1];
        if (ensure(*c->acts[l], (size_t)E2 * co * nxe)) return SFEM_FAILURE;
        if (cnn_conv_forward(L, E2, ci, co, in, c->W.p + c->offset[l], l > 0, c->acts[l]->p, s)) return SFEM_FAILURE;
        in = c->acts[l]->p;
    }
    return SFEM_SUCCESS;
}

int forward_f(sfem_amd_cnn *c, int L, ptrdiff_t E, const float *z, float *out, hipStream_t s) {
    const size_t n = (size_t)E * c->io * nxe_of(L);
    if (ensure(c->x2, 2 * n)) return SFEM_FAILURE;
    if (cnn_make_pm((ptrdiff_t)n, z, c->x2.p, s)) return SFEM_FAILURE;
    if (forward_g(c, L, 2 * E, s)) return SFEM_FAILURE;
    const float *g = c->acts[c->layers - 1]->p;
    return cnn_sub_halves((ptrdiff_t)n, g, g + n, out, s);
}

}  // namespace

namespace sfem_amd {
// Weights [layer][Cout][Cin][27] (hidden 16, 4 layers) -> BF16 in the lane order of the fused
// inference kernel: layer 1 [64 lanes][8], layers 2-3 [14 offset pairs][64][8], layer 4 [2][64][4].
void cnn_pack_weights(const float *W, uint16_t *packed) {
    const float *W1 = W, *W2 = W + 432, *W3 = W + 432 + 6912, *W4 = W + 432 + 2 * 6912;
    uint16_t *o = packed;
    for (int lane = 0; lane < 64; lane++)
        for (int j = 0; j < 8; j++) {
            const int co = lane % 16, k = 8 * (lane / 16) + j;
            *o++ = to_bf16(k < 27 ? W1[co * 27 + k] : 0.f);
        }
    for (const float *Wl : {W2, W3})
        for (int p = 0; p < 14; p++)
            for (int lane = 0; lane < 64; lane++)
                for (int j = 0; j < 8; j++) {
                    const int co = lane % 16, g = lane / 16, off = 2 * p + g / 2, ci = 8 * (g % 2) + j;
                    *o++ = to_bf16(off < 27 ? Wl[(co * 16 + ci) * 27 + off] : 0.f);
                }
    for (int h = 0; h < 2; h++)
        for (int lane = 0; lane < 64; lane++)
            for (int j = 0; j < 4; j++) {
                const int k = 16 * h + lane % 16, ci = 4 * (lane / 16) + j;
                *o++ = to_bf16(k < 27 ? W4[ci * 27 + k] : 0.f);
            }
}
}  // namespace sfem_amd

extern "C" {

int sfem_amd_cnn_create(int hidden, int layers, const float *weights, unsigned seed, sfem_amd_cnn **out) {
    return sfem_amd_cnn_create_channels(1, hidden, layers, weights, seed, out);
}

int sfem_amd_cnn_create_channels(int channels, int hidden, int layers, const float *weights, unsigned seed,
                                 sfem_amd_cnn **out) {
    SFEM_AMD_REQUIRE(hidden >= 1 && hidden <= 16 && layers >= 2 && layers <= 8, "cnn: hidden in 1..16, layers in 2..8");
    SFEM_AMD_REQUIRE(channels == 1 || channels == 3, "cnn: 1 or 3 input/output channels");
    auto c = std::make_unique<sfem_amd_cnn>();
    c->hidden = hidden;
    c->layers = layers;
    c->io = channels;
    c->chans.push_back(channels);
    for (int l = 0; l < layers - 1; l++) c->chans.push_back(hidden);
    c->chans.push_back(channels);
    size_t off = 0;
    for (int l = 0; l < layers; l++) {
        c->offset.push_back(off);
        off += (size_t)27 * c->chans[l] * c->chans[l + 1];
    }
    c->nparams = off;
    std::vector<float> h(off);
    if (weights) {
        std::copy(weights, weights + off, h.begin());
    } else {
        std::mt19937 rng(seed);
        for (int l = 0; l < layers; l++) {
            std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (27.f * c->chans[l])));
            for (size_t i = 0; i < (size_t)27 * c->chans[l] * c->chans[l + 1]; i++) h[c->offset[l] + i] = nd(rng);
        }
    }
    if (c->W.upload(h.data(), off) || c->grad.alloc(off) || c->m.alloc(off) || c->v.alloc(off)) return SFEM_FAILURE;
    SFEM_AMD_HIP_CHECK(hipMemset(c->m.p, 0, sizeof(float) * off));
    SFEM_AMD_HIP_CHECK(hipMemset(c->v.p, 0, sizeof(float) * off));
    for (int l = 0; l < layers; l++) c->acts.push_back(new DevArray<float>());
    if (c->ws.alloc(512)) return SFEM_FAILURE;
    *out = c.release();
    return SFEM_SUCCESS;
}

int sfem_amd_cnn_num_params(const sfem_amd_cnn *c) { return c ? (int)c->nparams : -1; }
int sfem_amd_cnn_channels(const sfem_amd_cnn *c) { return c ? c->io : -1; }

int sfem_amd_cnn_get_weights(const sfem_amd_cnn *c, float *w) {
    SFEM_AMD_REQUIRE(c && w, "cnn_get_weights: arguments");
    SFEM_AMD_HIP_CHECK(hipMemcpy(w, c->W.p, sizeof(float) * c->nparams, hipMemcpyDeviceToHost));
    return SFEM_SUCCESS;
}

int sfem_amd_cnn_set_weights(sfem_amd_cnn *c, const float *w) {
    SFEM_AMD_REQUIRE(c && w, "cnn_set_weights: arguments");
    SFEM_AMD_HIP_CHECK(hipMemcpy(c->W.p, w, sizeof(float) * c->nparams, hipMemcpyHostToDevice));
    c->w16_valid = false;
    bump_state_epoch();
    return SFEM_SUCCESS;
}

void sfem_amd_cnn_destroy(sfem_amd_cnn *c) { delete c; }

int sfem_amd_cnn_forward(sfem_amd_cnn *c, int L, ptrdiff_t E, const float *z, float *out, int mode, void *stream) {
    SFEM_AMD_REQUIRE(c && z && out, "cnn_forward: arguments");
    if (check_level(L)) return SFEM_FAILURE;
    hipStream_t s = as_stream(stream);
    if (mode == 1) {
        SFEM_AMD_REQUIRE(c->hidden == 16 && c->layers == 4 && c->io == 1,
                         "cnn_forward: the MFMA path is built for the scalar network of 4 layers of 16");
        if (!c->w16_valid) {
            std::vector<float> h(c->nparams);
            SFEM_AMD_HIP_CHECK(hipMemcpy(h.data(), c->W.p, sizeof(float) * c->nparams, hipMemcpyDeviceToHost));
            std::vector<uint16_t> hb(kCnnPackedWeights);
            cnn_pack_weights(h.data(), hb.data());
            // allocated once: a graph captured earlier keeps a valid pointer (it re-captures anyway,
            // since weight changes move the state epoch)
            if (c->W16.n != hb.size() && c->W16.alloc(hb.size())) return SFEM_FAILURE;
            SFEM_AMD_HIP_CHECK(hipMemcpy(c->W16.p, hb.data(), sizeof(uint16_t) * hb.size(), hipMemcpyHostToDevice));
            c->w16_valid = true;
        }
        return cnn_infer_mfma(L, E, c->W16.p, z, out, s);
    }
    return forward_f(c, L, E, z, out, s);
}

int sfem_amd_cnn_loss_grad(sfem_amd_cnn *c, int L, ptrdiff_t E, const float *z, const float *target, double *loss,
                           float *grad_host, void *stream) {
    SFEM_AMD_REQUIRE(c && z && target && loss, "cnn_loss_grad: arguments");
    if (check_level(L)) return SFEM_FAILURE;
    hipStream_t s = as_stream(stream);
    const size_t nxe = nxe_of(L), n = (size_t)E * c->io * nxe;
    if (ensure(c->out2, n) || ensure(c->dpm, 2 * n)) return SFEM_FAILURE;
    if (forward_f(c, L, E, z, c->out2.p, s)) return SFEM_FAILURE;
    double err = 0, norm = 0;
    if (cnn_sq_err((ptrdiff_t)n, c->out2.p, target, c->ws.p, &err, &norm, s)) return SFEM_FAILURE;
    SFEM_AMD_REQUIRE(norm > 0, "cnn_loss_grad: zero target");
    *loss = err / norm;
    if (cnn_dloss((ptrdiff_t)n, c->out2.p, target, (float)(1.0 / norm), c->dpm.p, s)) return SFEM_FAILURE;
    SFEM_AMD_HIP_CHECK(hipMemsetAsync(c->grad.p, 0, sizeof(float) * c->nparams, s));
    const ptrdiff_t E2 = 2 * E;
    const int nblocks = (int)std::min<ptrdiff_t>(E2, 512);
    if (ensure(c->part, (size_t)nblocks * 27 * 16 * 16)) return SFEM_FAILURE;
    const float *d = c->dpm.p;
    for (int l = c->layers - 1; l >= 0; l--) {
        const int ci = c->chans[l], co = c->chans[l + 1];
        const float *in = l == 0 ? c->x2.p : c->acts[l - 1]->p;
        if (cnn_conv_backward_weights(L, E2, ci, co, d, in, l > 0, c->part.p, nblocks, c->grad.p + c->offset[l], s))
            return SFEM_FAILURE;
        if (l > 0) {
            DevArray<float> &dn = (l % 2) ? c->dA : c->dB;
            if (ensure(dn, (size_t)E2 * ci * nxe)) return SFEM_FAILURE;
            if (cnn_conv_backward_data(L, E2, ci, co, d, c->W.p + c->offset[l], c->acts[l - 1]->p, 1, dn.p, s))
                return SFEM_FAILURE;
            d = dn.p;
        }
    }
    if (grad_host) {
        SFEM_AMD_HIP_CHECK(hipMemcpyAsync(grad_host, c->grad.p, sizeof(float) * c->nparams, hipMemcpyDeviceToHost, s));
        SFEM_AMD_HIP_CHECK(hipStreamSynchronize(s));
    }
    return SFEM_SUCCESS;
}

int sfem_amd_cnn_adam_step(sfem_amd_cnn *c, double lr, void *stream) {
    SFEM_AMD_REQUIRE(c, "cnn_adam_step: null");
    c->step++;
    c->w16_valid = false;
    bump_state_epoch();
    return cnn_adam((int)c->nparams, c->grad.p, c->W.p, c->m.p, c->v.p, (float)lr, 0.9f, 0.999f, 1e-8f, c->step,
                    as_stream(stream));
}

int sfem_amd_gather_elements(const sfem_amd_mesh *mesh, int rt, int bs, int comp, const void *vec, const void *scale,
                             float *out, void *stream) {
    SFEM_AMD_REQUIRE(mesh && vec && out && mesh->type == SFEM_AMD_SSHEX8, "gather_elements: sshex8 mesh and vectors");
    SFEM_AMD_REQUIRE((comp >= 0 && comp < bs) || comp == -1, "gather_elements: component");
    const int c0 = comp < 0 ? 0 : comp, c1 = comp < 0 ? bs : comp + 1;
    for (int c = c0; c < c1; c++) {
        const int err = rt == SFEM_AMD_FLOAT64
                                ? gather_elements<double>(mesh->E, mesh->nv, mesh->idx.p, bs, c, (const double *)vec,
                                                          (const double *)scale, out, as_stream(stream), comp < 0)
                                : gather_elements<float>(mesh->E, mesh->nv, mesh->idx.p, bs, c, (const float *)vec,
                                                         (const float *)scale, out, as_stream(stream), comp < 0);
        if (err) return err;
    }
    return SFEM_SUCCESS;
}

}  // extern "C"

// ------------------------------------------------------------------ the smoother operator
namespace {

template <typename T>
struct LearnedSmootherOp final : OpBase {
    std::shared_ptr<OpBase> A;  // constrained level operator
    const sfem_amd_mesh *mesh{nullptr};
    sfem_amd_cnn *cnn{nullptr};
    DevArray<T> invdiag, r, corr;
    DevArray<float> z, out;
    int bs{1}, steps{1}, mode{0};
    bool constrained{false};
    double damping{1.0};
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return A->rows(); }
    int apply(const void *bv, void *xv, hipStream_t s) override {
        const ptrdiff_t n = rows();
        const T *b = (const T *)bv;
        T *x = (T *)xv;
        const int rt = real_type();
        for (int it = 0; it < steps; it++) {
            if (A->apply_overwrite(x, r.p, s)) return SFEM_FAILURE;
            if (blas_axpby<T>(n, T(1), b, T(-1), r.p, s)) return SFEM_FAILURE;  // r = b - A x
            if (blas_values<T>(n, T(0), corr.p, s)) return SFEM_FAILURE;
            if (cnn->io == bs && bs > 1) {
                // the vector network: all components as channels of one evaluation
                for (int c = 0; c < bs; c++)
                    if (gather_elements<T>(mesh->E, mesh->nv, mesh->idx.p, bs, c, r.p, invdiag.p, z.p, s, true))
                        return SFEM_FAILURE;
                if (sfem_amd_cnn_forward(cnn, mesh->level, mesh->E, z.p, out.p, mode, s)) return SFEM_FAILURE;
                for (int c = 0; c < bs; c++)
                    if (scatter_elements<T>(mesh->E, mesh->nv, mesh->idx.p, bs, c, mesh->count.p, 1.f, out.p, corr.p, s,
                                            true))
                        return SFEM_FAILURE;
            } else {
                // the scalar network on each component in turn
                for (int c = 0; c < bs; c++) {
                    if (gather_elements<T>(mesh->E, mesh->nv, mesh->idx.p, bs, c, r.p, invdiag.p, z.p, s))
                        return SFEM_FAILURE;
                    if (sfem_amd_cnn_forward(cnn, mesh->level, mesh->E, z.p, out.p, mode, s)) return SFEM_FAILURE;
                    if (scatter_elements<T>(mesh->E, mesh->nv, mesh->idx.p, bs, c, mesh->count.p, 1.f, out.p, corr.p, s))
                        return SFEM_FAILURE;
                }
            }
            if (constrained && constrained_set_value(A.get(), rt, 0.0, corr.p, s)) return SFEM_FAILURE;
            if (blas_axpby<T>(n, (T)damping, corr.p, T(1), x, s)) return SFEM_FAILURE;
        }
        return SFEM_SUCCESS;
    }
};

}  // namespace

extern "C" int sfem_amd_learned_smoother_create(sfem_amd_op *A, const sfem_amd_mesh *mesh, sfem_amd_cnn *cnn,
                                                const void *diag, int bs, int steps, double damping, int mode,
                                                sfem_amd_op **out) {
    SFEM_AMD_REQUIRE(A && mesh && cnn && diag && out, "learned_smoother: arguments");
    SFEM_AMD_REQUIRE(mesh->type == SFEM_AMD_SSHEX8, "learned_smoother: sshex8 level required");
    if (check_level(mesh->level)) return SFEM_FAILURE;
    SFEM_AMD_REQUIRE(A->impl->rows() == mesh->nnodes * bs, "learned_smoother: operator size != mesh dofs");
    SFEM_AMD_REQUIRE(cnn->io == 1 || cnn->io == bs, "learned_smoother: network channels must be 1 or the block size");
    SFEM_AMD_REQUIRE(cnn->io == 1 || mode == 0, "learned_smoother: the vector network runs the FP32 path (mode 0)");
    if (mesh_counts(mesh, nullptr)) return SFEM_FAILURE;
    auto make = [&](auto op) -> int {
        using T = typename std::remove_reference<decltype(*op)>::type;
        (void)sizeof(T);
        op->A = A->impl;
        op->mesh = mesh;
        op->cnn = cnn;
        op->bs = bs;
        op->steps = steps;
        op->damping = damping;
        op->mode = mode;
        op->constrained = is_constrained(A->impl.get());
        const ptrdiff_t n = A->impl->rows();
        const size_t ne = (size_t)mesh->E * mesh->nv * cnn->io;
        if (op->r.alloc(n) || op->corr.alloc(n) || op->invdiag.alloc(n) || op->z.alloc(ne) || op->out.alloc(ne))
            return SFEM_FAILURE;
        *out = wrap_op(op);
        return SFEM_SUCCESS;
    };
    const int rt = A->impl->real_type();
    int err;
    if (rt == SFEM_AMD_FLOAT64) {
        auto op = std::make_shared<LearnedSmootherOp<double>>();
        err = make(op);
        if (!err) err = blas_copy<double>(A->impl->rows(), (const double *)diag, op->invdiag.p, nullptr) ||
                        blas_reciprocal<double>(A->impl->rows(), 1.0, op->invdiag.p, nullptr);
    } else {
        auto op = std::make_shared<LearnedSmootherOp<float>>();
        err = make(op);
        if (!err) err = blas_copy<float>(A->impl->rows(), (const float *)diag, op->invdiag.p, nullptr) ||
                        blas_reciprocal<float>(A->impl->rows(), 1.f, op->invdiag.p, nullptr);
    }
    return err ? SFEM_FAILURE : SFEM_SUCCESS;
}
