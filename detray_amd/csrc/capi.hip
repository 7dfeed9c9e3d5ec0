// C ABI of detray_amd: device-resident detector / field handles and the
// batched propagation entry point (include/detray_amd.h).
//
// dtr_propagate launches one work-item per track running detray's
// propagate() loop (core/include/detray/propagator/propagator.hpp:262-400)
// with the actor chain pathlimit_aborter, parameter_updater
// <pointwise_material_interactor>, surface_sequencer.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstring>
#include <new>

#include "detray_amd.h"
#include "device/propagator.hpp"

struct dtr_detector {
    dtr::DevDetector dev;
    void* blocks[8];
    uint32_t n_volumes;
};

struct dtr_field {
    dtr::DevField dev;
    float* values;
};

namespace {

int to_status(hipError_t e) {
    if (e == hipSuccess) return DTR_SUCCESS;
    if (e == hipErrorOutOfMemory) return DTR_ERROR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return DTR_ERROR_NO_DEVICE;
    return DTR_ERROR_HIP;
}

template <typename T>
int upload(const T* host, uint32_t n, void** dst) {
    *dst = nullptr;
    if (n == 0u) return DTR_SUCCESS;
    if (!host) return DTR_ERROR_INVALID_ARGUMENT;
    hipError_t e = hipMalloc(dst, sizeof(T) * static_cast<size_t>(n));
    if (e != hipSuccess) return to_status(e);
    e = hipMemcpy(*dst, host, sizeof(T) * static_cast<size_t>(n), hipMemcpyHostToDevice);
    return to_status(e);
}

__global__ void __launch_bounds__(256) propagate_fp64_kernel(
    dtr::DevDetector det, dtr::DevField field, dtr::DevConfig cfg,
    const double* __restrict__ fp, const double* __restrict__ init_cov,
    const double* __restrict__ bound0, const uint32_t* __restrict__ start_sf,
    dtr::DevResults out, uint32_t n) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    dtr::propagate_track(det, field, cfg, fp, init_cov, bound0, start_sf, out, n, gid);
}

}  // namespace

extern "C" {

const char* dtr_version(void) { return "detray_amd 0.1.0 (gfx950)"; }

void dtr_default_config(dtr_prop_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    // stepping_config.hpp:27-48
    c->min_stepsize = 1e-4f * 1.f;
    c->rk_error_tol = 1e-4f * 1.f;
    c->step_constraint = FLT_MAX;
    c->path_limit = 5.f * 1000.f;
    c->max_rk_updates = 10000u;
    c->use_mean_loss = 1;
    c->use_eloss_gradient = 0;
    c->use_field_gradient = 0;
    c->do_covariance_transport = 1;
    // intersection_config.hpp:20-36
    c->min_mask_tolerance = 1e-5f * 1.f;
    c->max_mask_tolerance = 3.f * 1.f;
    c->mask_tolerance_scalor = 5e-2f;
    c->path_tolerance = 1.f * 1e-3f;
    c->overstep_tolerance = -1000.f * 1e-3f;
    // navigation_config.hpp:23-40
    c->search_window[0] = 0u;
    c->search_window[1] = 0u;
    c->accumulated_error = 0.001f;
    c->n_scattering_stddev = 3;
    c->estimate_scattering_noise = 1;
    // pointwise_material_interactor::state
    c->do_energy_loss = 1;
    c->do_multiple_scattering = 1;
    c->use_material = 1;
    c->max_surfaces = 64u;
}

int dtr_detector_create(const dtr_detector_desc* d, dtr_detector** out) {
    if (!d || !out) return DTR_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (d->n_volumes == 0u || d->n_surfaces == 0u) return DTR_ERROR_INVALID_ARGUMENT;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return DTR_ERROR_NO_DEVICE;
    dtr_detector* h = new (std::nothrow) dtr_detector{};
    if (!h) return DTR_ERROR_OUT_OF_MEMORY;
    int rc = DTR_SUCCESS;
    if (!rc) rc = upload(d->volumes, d->n_volumes, &h->blocks[0]);
    if (!rc) rc = upload(d->surfaces, d->n_surfaces, &h->blocks[1]);
    if (!rc) rc = upload(d->transforms, d->n_transforms, &h->blocks[2]);
    if (!rc) rc = upload(d->masks, d->n_masks, &h->blocks[3]);
    if (!rc) rc = upload(d->materials, d->n_materials, &h->blocks[4]);
    if (!rc) rc = upload(d->grids, d->n_grids, &h->blocks[5]);
    if (!rc) rc = upload(d->bf_surfaces, d->n_bf, &h->blocks[6]);
    if (!rc) rc = upload(d->grid_bins, d->n_bins, &h->blocks[7]);
    if (rc) {
        dtr_detector_destroy(h);
        return rc;
    }
    h->dev.volumes = static_cast<const dtr_volume*>(h->blocks[0]);
    h->dev.surfaces = static_cast<const dtr_surface*>(h->blocks[1]);
    h->dev.transforms = static_cast<const dtr_transform*>(h->blocks[2]);
    h->dev.masks = static_cast<const dtr_mask*>(h->blocks[3]);
    h->dev.materials = static_cast<const dtr_material_slab*>(h->blocks[4]);
    h->dev.grids = static_cast<const dtr_grid*>(h->blocks[5]);
    h->dev.bf = static_cast<const uint32_t*>(h->blocks[6]);
    h->dev.bins = static_cast<const uint32_t*>(h->blocks[7]);
    h->n_volumes = d->n_volumes;
    *out = h;
    return DTR_SUCCESS;
}

int dtr_detector_destroy(dtr_detector* det) {
    if (!det) return DTR_ERROR_INVALID_ARGUMENT;
    int rc = DTR_SUCCESS;
    for (void* p : det->blocks) {
        if (p && hipFree(p) != hipSuccess) rc = DTR_ERROR_HIP;
    }
    delete det;
    return rc;
}

int dtr_field_create(const dtr_field_desc* d, dtr_field** out) {
    if (!d || !out) return DTR_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (d->kind != DTR_FIELD_CONSTANT && d->kind != DTR_FIELD_GRID) return DTR_ERROR_INVALID_ARGUMENT;
    dtr_field* h = new (std::nothrow) dtr_field{};
    if (!h) return DTR_ERROR_OUT_OF_MEMORY;
    h->dev.kind = d->kind;
    for (int i = 0; i < 3; ++i) {
        h->dev.b[i] = d->b[i];
        h->dev.n[i] = d->n[i];
        h->dev.min[i] = d->min[i];
        h->dev.pitch[i] = d->pitch[i];
    }
    if (d->kind == DTR_FIELD_GRID) {
        if (!d->values || d->n[0] < 2u || d->n[1] < 2u || d->n[2] < 2u || !(d->pitch[0] > 0.0) ||
            !(d->pitch[1] > 0.0) || !(d->pitch[2] > 0.0)) {
            delete h;
            return DTR_ERROR_INVALID_ARGUMENT;
        }
        const size_t count = 3ull * d->n[0] * d->n[1] * d->n[2];
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, count * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(p, d->values, count * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (p) (void)hipFree(p);
            delete h;
            return to_status(e);
        }
        h->values = static_cast<float*>(p);
    }
    h->dev.values = h->values;
    *out = h;
    return DTR_SUCCESS;
}

int dtr_field_destroy(dtr_field* f) {
    if (!f) return DTR_ERROR_INVALID_ARGUMENT;
    int rc = DTR_SUCCESS;
    if (f->values && hipFree(f->values) != hipSuccess) rc = DTR_ERROR_HIP;
    delete f;
    return rc;
}

int dtr_propagate(const dtr_detector* det, const dtr_field* field, const dtr_prop_config* c,
                  const dtr_track_batch* in, dtr_result_batch* out, void* stream) {
    if (!det || !field || !c || !in || !out) return DTR_ERROR_INVALID_ARGUMENT;
    if (in->n == 0u) return DTR_SUCCESS;
    if (!in->free_params && !in->bound_params) return DTR_ERROR_INVALID_ARGUMENT;
    if (in->bound_params && !in->start_surface) return DTR_ERROR_INVALID_ARGUMENT;
    if (c->max_surfaces != 0u && (!out->surfaces || !out->n_surfaces)) return DTR_ERROR_INVALID_ARGUMENT;
    if (c->use_field_gradient || c->use_eloss_gradient) return DTR_ERROR_UNSUPPORTED;
    if (c->max_rk_updates == 0u) return DTR_ERROR_INVALID_ARGUMENT;
    dtr::DevConfig dc{};
    dc.min_stepsize = c->min_stepsize;
    dc.rk_error_tol = c->rk_error_tol;
    dc.step_constraint = c->step_constraint;
    dc.path_limit = c->path_limit;
    dc.max_rk_updates = c->max_rk_updates;
    dc.min_mask_tol = c->min_mask_tolerance;
    dc.max_mask_tol = c->max_mask_tolerance;
    dc.mask_scalor = c->mask_tolerance_scalor;
    dc.path_tol = c->path_tolerance;
    dc.overstep_tol = c->overstep_tolerance;
    dc.win0 = c->search_window[0];
    dc.win1 = c->search_window[1];
    dc.accumulated_error = c->accumulated_error;
    dc.n_stddev = static_cast<double>(c->n_scattering_stddev);
    dc.max_surfaces = c->max_surfaces;
    dc.do_cov = c->do_covariance_transport;
    dc.estimate_noise = c->estimate_scattering_noise;
    dc.do_eloss = c->do_energy_loss;
    dc.do_ms = c->do_multiple_scattering;
    dc.use_material = c->use_material;
    dtr::DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                      c->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                      out->n_trials, out->status};
    const uint32_t block = 256;
    const uint32_t grid = (in->n + block - 1) / block;
    hipLaunchKernelGGL(propagate_fp64_kernel, dim3(grid), dim3(block), 0,
                       static_cast<hipStream_t>(stream), det->dev, field->dev, dc,
                       in->free_params, in->initial_cov, in->bound_params, in->start_surface,
                       r, in->n);
    return to_status(hipGetLastError());
}

}  // extern "C"
