// libtorch binding of the UMAP cross-entropy loss: a torch::autograd::Function
// whose forward and backward call the HIP kernels of libumappp_amd through
// its C ABI (umx_loss_forward / umx_loss_backward, include/umappp_amd.h).
// Gradients flow to the float32 embedding and to the curve parameters a and
// b (float64 scalars).  The backward pass recomputes the negative samples
// from the counter sampler (seed, step, edge key, draw) instead of storing
// them.  No reference counterpart: the reference only applies per-edge SGD
// steps (optimize_layout.hpp:145-176), whose directions this loss's gradient
// reproduces.
//
// Built with torch.utils.cpp_extension (a C++ extension, no device code of
// its own); the HIP stream is passed in from Python as an integer handle.
#include <torch/extension.h>

#include <string>

#include "umappp_amd.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void check(int rc) {
    if (rc != UMX_OK) throw std::runtime_error(std::string("umappp_amd loss: ") + umx_last_error());
}

void check_input(const Tensor& t, torch::ScalarType dt, const char* what) {
    TORCH_CHECK(t.is_cuda(), what, " must be a GPU tensor");
    TORCH_CHECK(t.scalar_type() == dt, what, " has the wrong dtype");
    TORCH_CHECK(t.is_contiguous(), what, " must be contiguous");
}

template <typename T>
const T* opt_ptr(const Tensor& t) {
    return t.defined() && t.numel() > 0 ? t.data_ptr<T>() : nullptr;
}

class UmapLoss : public torch::autograd::Function<UmapLoss> {
public:
    // emb: float32 (n, dim); a, b: float64 scalars; heads, tails: int32 (nnz);
    // keys: int64 (nnz) or empty; coef: float32 (nnz) or empty; num_neg:
    // int32 (nnz) or empty.  value = false skips the loss kernel (the result
    // is 0; the backward pass, which recomputes everything it needs, is
    // unchanged) for callers that only want the gradient.
    static Tensor forward(AutogradContext* ctx, Tensor emb, Tensor a, Tensor b, Tensor heads, Tensor tails,
                          Tensor keys, Tensor coef, Tensor num_neg, int64_t m, double gamma, int64_t seed,
                          int64_t step, double clip, int64_t stream, bool value) {
        check_input(emb, torch::kFloat32, "embedding");
        TORCH_CHECK(emb.dim() == 2, "embedding must be (n, dim)");
        check_input(heads, torch::kInt32, "heads");
        check_input(tails, torch::kInt32, "tails");
        TORCH_CHECK(heads.numel() == tails.numel(), "heads and tails differ in length");
        if (keys.numel()) check_input(keys, torch::kInt64, "keys");
        if (coef.numel()) check_input(coef, torch::kFloat32, "coef");
        if (num_neg.numel()) check_input(num_neg, torch::kInt32, "num_neg");
        Tensor ab = torch::stack({a.reshape({}), b.reshape({})}).to(emb.device(), torch::kFloat64).contiguous();
        Tensor loss = value ? torch::empty({}, emb.options().dtype(torch::kFloat64))
                            : torch::zeros({}, emb.options().dtype(torch::kFloat64));
        if (value)
            check(umx_loss_forward(emb.size(0), static_cast<int>(emb.size(1)), emb.data_ptr<float>(), heads.numel(),
                               heads.data_ptr<int32_t>(), tails.data_ptr<int32_t>(), opt_ptr<int64_t>(keys),
                               opt_ptr<float>(coef), opt_ptr<int32_t>(num_neg), static_cast<int>(m),
                               ab.data_ptr<double>(), gamma, static_cast<uint64_t>(seed), static_cast<int>(step),
                               loss.data_ptr<double>(), reinterpret_cast<void*>(stream)));
        ctx->save_for_backward({emb, ab, heads, tails, keys, coef, num_neg, a, b});
        ctx->saved_data["m"] = m;
        ctx->saved_data["gamma"] = gamma;
        ctx->saved_data["seed"] = seed;
        ctx->saved_data["step"] = step;
        ctx->saved_data["clip"] = clip;
        ctx->saved_data["stream"] = stream;
        return loss;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grads) {
        auto sv = ctx->get_saved_variables();
        Tensor emb = sv[0], ab = sv[1], heads = sv[2], tails = sv[3], keys = sv[4], coef = sv[5], num_neg = sv[6];
        Tensor a = sv[7], b = sv[8];
        Tensor go = grads[0].to(emb.device(), torch::kFloat64).contiguous();
        Tensor grad_emb = torch::empty_like(emb);
        Tensor grad_ab = torch::empty({2}, emb.options().dtype(torch::kFloat64));
        check(umx_loss_backward(emb.size(0), static_cast<int>(emb.size(1)), emb.data_ptr<float>(), heads.numel(),
                                heads.data_ptr<int32_t>(), tails.data_ptr<int32_t>(), opt_ptr<int64_t>(keys),
                                opt_ptr<float>(coef), opt_ptr<int32_t>(num_neg),
                                static_cast<int>(ctx->saved_data["m"].toInt()), ab.data_ptr<double>(),
                                ctx->saved_data["gamma"].toDouble(),
                                static_cast<uint64_t>(ctx->saved_data["seed"].toInt()),
                                static_cast<int>(ctx->saved_data["step"].toInt()),
                                ctx->saved_data["clip"].toDouble(), go.data_ptr<double>(),
                                grad_emb.data_ptr<float>(), grad_ab.data_ptr<double>(),
                                reinterpret_cast<void*>(ctx->saved_data["stream"].toInt())));
        return {grad_emb, grad_ab[0].to(a.options()).reshape(a.sizes()), grad_ab[1].to(b.options()).reshape(b.sizes()),
                Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(),
                Tensor(), Tensor()};
    }
};

Tensor umap_loss(Tensor emb, Tensor a, Tensor b, Tensor heads, Tensor tails, Tensor keys, Tensor coef,
                 Tensor num_neg, int64_t m, double gamma, int64_t seed, int64_t step, double clip, int64_t stream,
                 bool value) {
    return UmapLoss::apply(emb, a, b, heads, tails, keys, coef, num_neg, m, gamma, seed, step, clip, stream, value);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, mod) {
    mod.doc() = "UMAP cross-entropy loss with hand-written HIP forward/backward (torch::autograd::Function)";
    mod.def("umap_loss", &umap_loss, "UMAP cross-entropy loss of an embedding over an edge list");
}
