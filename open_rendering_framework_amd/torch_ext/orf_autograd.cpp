// orf_autograd.cpp -- libtorch autograd Function over the differentiable C-ABI.
//
// forward : orf_diff_render_ex   (textures, scattering coefficients -> per-view images)
// backward: orf_diff_adjoint_ex  (dLoss/dImages -> dLoss/dTextures, dLoss/dCoefficients, path replay)
// Both run on the current HIP stream of the textures' device; the render
// context, texture mapping and views live in a Setup owned by Python.
#include <torch/extension.h>

#include <c10/hip/HIPStream.h>

#include <memory>
#include <vector>

#include "orf_rt.h"

namespace {

struct Setup {
    orf_context ctx = nullptr;
    orf_diff_desc desc{};
    std::vector<orf_render_item> views;
    uint32_t frame_shift = 0x40000000u;
};

void check(orf_status s)
{
    TORCH_CHECK(s == ORF_SUCCESS, "orf_rt error ", (int)s, ": ", orf_last_error());
}

void check_tex(const torch::Tensor& t, int64_t n, const char* what)
{
    TORCH_CHECK(t.device().is_cuda(), what, " must be a device tensor");
    TORCH_CHECK(t.scalar_type() == torch::kFloat32, what, " must be float32");
    TORCH_CHECK(t.is_contiguous(), what, " must be contiguous");
    TORCH_CHECK(t.numel() == n, what, " has ", t.numel(), " elements, expected ", n);
}

class OrfRender : public torch::autograd::Function<OrfRender> {
public:
    static torch::Tensor forward(torch::autograd::AutogradContext* actx, torch::Tensor albedo, torch::Tensor rough,
                                 torch::Tensor scattering, int64_t setup)
    {
        auto* s = reinterpret_cast<Setup*>(setup);
        int64_t texels = (int64_t)s->desc.tex_w * s->desc.tex_h;
        check_tex(albedo, texels * 3, "albedo");
        check_tex(rough, texels, "roughness");
        const bool sss = s->desc.scattering_material >= 0;
        if (sss) check_tex(scattering, 6, "scattering coefficients");
        int64_t nv = (int64_t)s->views.size();
        auto img = torch::empty({nv, s->views[0].height, s->views[0].width, 3}, albedo.options());
        hipStream_t st = c10::hip::getCurrentHIPStream(albedo.device().index()).stream();
        check(orf_diff_render_ex(s->ctx, &s->desc, s->views.data(), (int32_t)nv, albedo.data_ptr<float>(),
                                 rough.data_ptr<float>(), sss ? scattering.data_ptr<float>() : nullptr,
                                 img.data_ptr<float>(), st));
        actx->save_for_backward({albedo, rough, sss ? scattering : torch::zeros({6}, albedo.options())});
        actx->saved_data["setup"] = setup;
        return img;
    }

    static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* actx,
                                                 torch::autograd::tensor_list grads)
    {
        auto saved = actx->get_saved_variables();
        auto albedo = saved[0], rough = saved[1], scattering = saved[2];
        auto* s = reinterpret_cast<Setup*>(actx->saved_data["setup"].toInt());
        const bool sss = s->desc.scattering_material >= 0;
        auto gi = grads[0].contiguous().to(torch::kFloat32);
        auto ga = torch::zeros_like(albedo), gr = torch::zeros_like(rough), gs = torch::zeros_like(scattering);
        hipStream_t st = c10::hip::getCurrentHIPStream(albedo.device().index()).stream();
        check(orf_diff_adjoint_ex(s->ctx, &s->desc, s->views.data(), (int32_t)s->views.size(),
                                  albedo.data_ptr<float>(), rough.data_ptr<float>(),
                                  sss ? scattering.data_ptr<float>() : nullptr, gi.data_ptr<float>(),
                                  ga.data_ptr<float>(), gr.data_ptr<float>(), sss ? gs.data_ptr<float>() : nullptr,
                                  s->frame_shift, st));
        return {ga, gr, sss ? gs : torch::Tensor(), torch::Tensor()};
    }
};

// desc: [albedo_material, albedo_uv, ax0, az0, ax1, az1, rough_material, rough_uv, rx0, ry0, rz0, r3, tex_w, tex_h,
//        scattering_material (optional, -1: none)]
// views: [n, 16] = width, height, spp, frame0, eye[3], U[3], V[3], W[3]
int64_t make_setup(int64_t ctx, std::vector<double> desc, std::vector<std::vector<double>> views, int64_t frame_shift)
{
    TORCH_CHECK(desc.size() == 14 || desc.size() == 15, "desc must have 14 or 15 entries");
    TORCH_CHECK(!views.empty(), "at least one view");
    auto s = std::make_unique<Setup>();
    s->ctx = reinterpret_cast<orf_context>(ctx);
    s->desc.albedo_material = (int32_t)desc[0];
    s->desc.albedo_uv = (int32_t)desc[1];
    for (int k = 0; k < 4; ++k) s->desc.albedo_uv_params[k] = (float)desc[2 + k];
    s->desc.roughness_material = (int32_t)desc[6];
    s->desc.roughness_uv = (int32_t)desc[7];
    for (int k = 0; k < 4; ++k) s->desc.roughness_uv_params[k] = (float)desc[8 + k];
    s->desc.tex_w = (int32_t)desc[12];
    s->desc.tex_h = (int32_t)desc[13];
    s->desc.scattering_material = desc.size() == 15 ? (int32_t)desc[14] : -1;
    for (auto& v : views) {
        TORCH_CHECK(v.size() == 16, "a view has 16 entries");
        orf_render_item it{};
        it.width = (int32_t)v[0];
        it.height = (int32_t)v[1];
        it.spp = (uint32_t)v[2];
        it.frame0 = (uint32_t)v[3];
        for (int k = 0; k < 3; ++k) {
            it.eye[k] = (float)v[4 + k];
            it.U[k] = (float)v[7 + k];
            it.V[k] = (float)v[10 + k];
            it.W[k] = (float)v[13 + k];
        }
        s->views.push_back(it);
    }
    s->frame_shift = (uint32_t)frame_shift;
    return reinterpret_cast<int64_t>(s.release());
}

void free_setup(int64_t setup) { delete reinterpret_cast<Setup*>(setup); }

torch::Tensor render(torch::Tensor albedo, torch::Tensor rough, int64_t setup)
{
    // no scattering parameter: an empty tensor on the textures' device stands in for it
    return OrfRender::apply(albedo, rough, torch::zeros({0}, albedo.options()), setup);
}

torch::Tensor render_sss(torch::Tensor albedo, torch::Tensor rough, torch::Tensor scattering, int64_t setup)
{
    return OrfRender::apply(albedo, rough, scattering, setup);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "libtorch autograd Function over the differentiable render C-ABI";
    m.def("make_setup", &make_setup);
    m.def("free_setup", &free_setup);
    m.def("render", &render);
    m.def("render_sss", &render_sss);
}
