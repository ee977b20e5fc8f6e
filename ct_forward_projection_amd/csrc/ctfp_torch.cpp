// Differentiable polyenergetic projection for PyTorch: a torch::autograd::Function
// over the C ABI of libctfp.  Forward lays the density tensor out on the
// current HIP stream and projects it (ctfp_forward), keeping every ray's line
// integrals L (n_mat floats per ray) unless asked not to; backward runs the
// hand-written adjoint (ctfp_backward) from those L, so it neither marches
// the volume again nor needs the densities laid out.  Without kept L,
// backward lays the saved densities out again and the adjoint re-marches.
//
// The reference has no gradient; forward is its per-view projection call
// (CBCTSinMatPolyForwardProjGrid::computePolyForwProj,
// Source/CBCTSinMatPolyForwardProjGrid.cpp:82-105, through
// forwardSinMatProjGridGPU in Source/CUDAFun/CBCTPolyProjGPU.cu).
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include <cstdint>
#include <vector>

#include "ctfp.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

ctfp_context *as_context(int64_t handle)
{
    TORCH_CHECK(handle != 0, "ctfp: null context handle");
    return reinterpret_cast<ctfp_context *>(static_cast<intptr_t>(handle));
}

void check(int rc, const char *what)
{
    TORCH_CHECK(rc == CTFP_OK, what, " failed (status ", rc, "): ", ctfp_last_error());
}

struct Scan {
    ctfp_scan_params s;
    int device;
};

Scan scan_of(ctfp_context *c)
{
    Scan r{};
    check(ctfp_get_scan(c, &r.s), "ctfp_get_scan");
    check(ctfp_get_device(c, &r.device), "ctfp_get_device");
    return r;
}

// rho: [n_mat][p_num_z][p_num_y][p_num_x] float32, contiguous, on the context's device.
void check_densities(const torch::Tensor &rho, const Scan &sc)
{
    TORCH_CHECK(rho.is_cuda(), "ctfp: densities must be on the GPU");
    TORCH_CHECK(rho.scalar_type() == torch::kFloat32, "ctfp: densities must be float32");
    TORCH_CHECK(rho.is_contiguous(), "ctfp: densities must be contiguous");
    TORCH_CHECK(rho.dim() == 4 && rho.size(0) >= 1 && rho.size(0) <= CTFP_MAX_MATERIALS,
                "ctfp: densities must be [n_mat (1..", CTFP_MAX_MATERIALS, ")][nz][ny][nx]");
    TORCH_CHECK(rho.size(1) == sc.s.p_num_z && rho.size(2) == sc.s.p_num_y &&
                    rho.size(3) == sc.s.p_num_x,
                "ctfp: density volume ", rho.sizes(), " does not match the scan's phantom (",
                sc.s.p_num_z, ", ", sc.s.p_num_y, ", ", sc.s.p_num_x, ")");
    TORCH_CHECK(rho.get_device() == sc.device, "ctfp: densities are on cuda:", rho.get_device(),
                ", the context on cuda:", sc.device);
}

void load(ctfp_context *c, const torch::Tensor &rho, int64_t storage_dtype, void *stream)
{
    const int64_t nvox = rho.size(1) * rho.size(2) * rho.size(3);
    const float *ptrs[CTFP_MAX_MATERIALS] = {};
    for (int64_t m = 0; m < rho.size(0); m++) ptrs[m] = rho.data_ptr<float>() + m * nvox;
    check(ctfp_set_volumes_async(c, (int)rho.size(0), ptrs, (int)storage_dtype, stream),
          "ctfp_set_volumes_async");
}

const float *grad_ptr(torch::Tensor &g, int64_t vc, const ctfp_scan_params &s)
{
    if (!g.defined()) return nullptr;
    TORCH_CHECK(g.numel() == vc * s.d_num_v * s.d_num_u, "ctfp: upstream gradient of wrong size");
    g = g.to(torch::kFloat32).contiguous();
    return g.data_ptr<float>();
}

struct Project : public torch::autograd::Function<Project> {
    static variable_list forward(AutogradContext *actx, torch::Tensor rho, int64_t handle,
                                 int64_t view_begin, int64_t view_count, int64_t storage_dtype,
                                 bool keep_line_integrals)
    {
        ctfp_context *c = as_context(handle);
        const Scan sc = scan_of(c);
        check_densities(rho, sc);
        if (view_count < 0) view_count = sc.s.proj_num - view_begin;
        TORCH_CHECK(view_begin >= 0 && view_count >= 1 && view_begin + view_count <= sc.s.proj_num,
                    "ctfp: view range [", view_begin, ", ", view_begin + view_count,
                    ") outside [0, ", sc.s.proj_num, ")");
        void *stream = c10::hip::getCurrentHIPStream(rho.get_device()).stream();
        load(c, rho, storage_dtype, stream);
        auto I = torch::empty({view_count, sc.s.d_num_v, sc.s.d_num_u}, rho.options());
        auto IA = torch::empty_like(I);
        // [view][m][v][u], as ctfp_backward reads them
        torch::Tensor L;
        if (keep_line_integrals)
            L = torch::empty({view_count, rho.size(0), sc.s.d_num_v, sc.s.d_num_u}, rho.options());
        check(ctfp_forward(c, view_begin, view_count, I.data_ptr<float>(), IA.data_ptr<float>(),
                           L.defined() ? L.data_ptr<float>() : nullptr, stream),
              "ctfp_forward");
        actx->set_materialize_grads(false); // an unused output passes no gradient
        actx->save_for_backward({rho, L});
        actx->saved_data["handle"] = handle;
        actx->saved_data["view_begin"] = view_begin;
        actx->saved_data["view_count"] = view_count;
        actx->saved_data["storage_dtype"] = storage_dtype;
        return {I, IA};
    }

    static variable_list backward(AutogradContext *actx, variable_list grads)
    {
        const variable_list saved = actx->get_saved_variables();
        const torch::Tensor rho = saved[0], L = saved[1];
        const int64_t handle = actx->saved_data["handle"].toInt();
        const int64_t vb = actx->saved_data["view_begin"].toInt();
        const int64_t vc = actx->saved_data["view_count"].toInt();
        const int64_t dtype = actx->saved_data["storage_dtype"].toInt();
        ctfp_context *c = as_context(handle);
        const Scan sc = scan_of(c);
        torch::Tensor gI = grads[0], gIA = grads[1];
        const float *pI = grad_ptr(gI, vc, sc.s), *pIA = grad_ptr(gIA, vc, sc.s);
        auto grad = torch::zeros_like(rho);
        if (pI || pIA) {
            void *stream = c10::hip::getCurrentHIPStream(rho.get_device()).stream();
            // The context may hold other densities since forward ran.  With
            // kept L only the material count (the layout of L and of the
            // adjoint's channel groups) has to match; without, the adjoint
            // re-marches and needs these densities.
            int held_mat = 0, held_dtype = 0;
            if (!L.defined() || ctfp_get_volumes(c, &held_mat, &held_dtype) != CTFP_OK ||
                held_mat != rho.size(0))
                load(c, rho, dtype, stream);
            check(ctfp_backward(c, vb, vc, pI, pIA, L.defined() ? L.data_ptr<float>() : nullptr,
                                grad.data_ptr<float>(), stream),
                  "ctfp_backward");
        }
        return {grad, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
    }
};

std::vector<torch::Tensor> project(torch::Tensor rho, int64_t handle, int64_t view_begin,
                                   int64_t view_count, int64_t storage_dtype, bool keep_line_integrals)
{
    variable_list out = Project::apply(rho, handle, view_begin, view_count, storage_dtype, keep_line_integrals);
    return {out[0], out[1]};
}

} // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "Differentiable polyenergetic cone-beam projection (libctfp adjoint)";
    m.def("project", &project,
          "I, IAbsorb = project(rho, context, view_begin, view_count, storage_dtype, keep_line_integrals)",
          pybind11::arg("rho"), pybind11::arg("context"), pybind11::arg("view_begin") = 0,
          pybind11::arg("view_count") = -1, pybind11::arg("storage_dtype") = CTFP_F32,
          pybind11::arg("keep_line_integrals") = true);
}
