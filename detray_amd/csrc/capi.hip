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
#include <mutex>
#include <new>

#include <vector>

#include "detray_amd.h"
#include "host/tables.hpp"
#include "device/coop_nav.hpp"
#include "launch.hpp"

namespace {

using dtr::to_status;

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

// ---------------------------------------------------------------- schedule --
// Counting sort of the tracks by direction cell.
__device__ inline uint32_t direction_cell(const double* __restrict__ fp, uint64_t n, uint64_t i,
                                          uint32_t eta_bins_per_unit, uint32_t phi_bins, float eta_max) {
    const double dx = fp[4 * n + i], dy = fp[5 * n + i];
    double dz = fp[6 * n + i];
    dz = dz < -1.0 + 1e-12 ? -1.0 + 1e-12 : (dz > 1.0 - 1e-12 ? 1.0 - 1e-12 : dz);
    float eta = static_cast<float>(0.5 * log((1.0 + dz) / (1.0 - dz)));
    eta = eta < -eta_max ? -eta_max : (eta > eta_max ? eta_max : eta);
    const uint32_t n_eta = static_cast<uint32_t>(2.f * eta_max * static_cast<float>(eta_bins_per_unit)) + 1u;
    uint32_t ie = static_cast<uint32_t>((eta + eta_max) * static_cast<float>(eta_bins_per_unit));
    ie = ie >= n_eta ? n_eta - 1u : ie;
    const float phi = static_cast<float>(atan2(dy, dx));
    uint32_t ip = static_cast<uint32_t>((phi + 3.14159265f) * (static_cast<float>(phi_bins) / 6.2831853f));
    ip = ip >= phi_bins ? phi_bins - 1u : ip;
    return ie * phi_bins + ip;
}

__global__ void sort_clear_kernel(uint32_t* __restrict__ count, uint32_t cells) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cells) count[i] = 0u;
}

__global__ void sort_count_kernel(const double* __restrict__ fp, uint32_t n, uint32_t* __restrict__ count,
                                  uint32_t eta_bins_per_unit, uint32_t phi_bins, float eta_max) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    atomicAdd(&count[direction_cell(fp, n, i, eta_bins_per_unit, phi_bins, eta_max)], 1u);
}

// Exclusive scan of the cell counts in place, in two levels: blocks of 1024
// threads scan chunks of kScanChunk cells and write the chunk totals, one
// block scans the totals, and the chunk offsets are added back.
constexpr uint32_t kScanItems = 8u;
constexpr uint32_t kScanChunk = 1024u * kScanItems;

__global__ void __launch_bounds__(1024) sort_scan_chunk_kernel(uint32_t* __restrict__ count, uint32_t cells,
                                                               uint32_t* __restrict__ total) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanChunk + static_cast<uint64_t>(t) * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0u;
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < cells) ? count[base + i] : 0u;
        sum += v[i];
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1u; d < 1024u; d <<= 1) {
        const uint32_t x = (t >= d) ? part[t - d] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i) {
        if (base + i < cells) count[base + i] = run;
        run += v[i];
    }
    if (t == 1023u) total[blockIdx.x] = part[1023];
}

__global__ void sort_scan_add_kernel(uint32_t* __restrict__ count, uint32_t cells,
                                     const uint32_t* __restrict__ offset) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cells) count[i] += offset[i / kScanChunk];
}

// exclusive scan in place of a short array, one block of 1024 threads
__global__ void __launch_bounds__(1024) sort_scan_kernel(uint32_t* __restrict__ count, uint32_t cells) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (cells + 1023u) / 1024u;
    const uint32_t b = t * per, e = (b + per < cells) ? b + per : cells;
    uint32_t sum = 0u;
    for (uint32_t i = b; i < e; ++i) sum += count[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1u; d < 1024u; d <<= 1) {
        const uint32_t v = (t >= d) ? part[t - d] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t c = count[i];
        count[i] = run;
        run += c;
    }
}

__global__ void sort_scatter_kernel(const double* __restrict__ fp, uint32_t n, uint32_t* __restrict__ offset,
                                    uint32_t* __restrict__ order, uint32_t eta_bins_per_unit, uint32_t phi_bins,
                                    float eta_max) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t cell = direction_cell(fp, n, i, eta_bins_per_unit, phi_bins, eta_max);
    order[atomicAdd(&offset[cell], 1u)] = i;
}

// cells of the sort (at most 2^26; the configuration limits keep the
// product in range)
uint32_t sort_cells(const dtr_sort_config* c) {
    const uint64_t n_eta = static_cast<uint64_t>(2.f * c->eta_max * static_cast<float>(c->eta_bins_per_unit)) + 1u;
    const uint64_t cells = n_eta * c->phi_bins;
    return cells > (1ull << 26) ? 0u : static_cast<uint32_t>(cells);
}

// words of workspace: the cell counts, then the chunk totals of the scan
size_t sort_words(uint32_t cells) {
    return static_cast<size_t>(cells) + (cells + kScanChunk - 1u) / kScanChunk;
}

// Workspace of dtr_sort_tracks_static: a file-scope device array holding the
// cell counts and chunk totals of up to kStaticSortCells cells (4 MiB), one
// copy per GPU for the life of the process.  In host code `scratch` is only
// a stand-in for the symbol: its device address is taken once per device
// with hipGetSymbolAddress and used wherever the host hands the memory to a
// kernel.
constexpr uint32_t kStaticSortCells = 1u << 20;
constexpr size_t kStaticSortWords = static_cast<size_t>(kStaticSortCells) + kStaticSortCells / kScanChunk;
static __device__ __align__(256) uint32_t scratch[kStaticSortWords];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
uint32_t* scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (set up on the first
// call on that device).
int static_scratch(uint32_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return to_status(e);
    if (dev < 0 || dev >= kMaxDevices) return DTR_ERROR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return to_status(e);
        if (!p) return DTR_ERROR_HIP;
        scratch_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = scratch_ptr[dev];
    return DTR_SUCCESS;
}

int sort_config_ok(const dtr_sort_config* c) {
    return c && c->eta_bins_per_unit != 0u && c->phi_bins != 0u && c->eta_max > 0.f && c->eta_max <= 20.f &&
           c->eta_bins_per_unit <= 1000u && c->phi_bins <= 100000u;
}

// the counting sort's kernels on `stream` with `workspace` (sort_words(cells)
// words): clear, count, two-level exclusive scan, scatter
int sort_launch(const dtr_sort_config* c, const double* fp, uint32_t n, uint32_t* order, uint32_t* workspace,
                uint32_t cells, hipStream_t st) {
    uint32_t* count = workspace;
    uint32_t* total = count + cells;
    const uint32_t chunks = static_cast<uint32_t>((cells + kScanChunk - 1u) / kScanChunk);
    hipLaunchKernelGGL(sort_clear_kernel, dim3((cells + 255u) / 256u), dim3(256), 0, st, count, cells);
    hipLaunchKernelGGL(sort_count_kernel, dim3((n + 255u) / 256u), dim3(256), 0, st, fp, n, count,
                       c->eta_bins_per_unit, c->phi_bins, c->eta_max);
    hipLaunchKernelGGL(sort_scan_chunk_kernel, dim3(chunks), dim3(1024), 0, st, count, cells, total);
    if (chunks > 1u) {
        hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(1024), 0, st, total, chunks);
        hipLaunchKernelGGL(sort_scan_add_kernel, dim3((cells + 255u) / 256u), dim3(256), 0, st, count, cells, total);
    }
    hipLaunchKernelGGL(sort_scatter_kernel, dim3((n + 255u) / 256u), dim3(256), 0, st, fp, n, count, order,
                       c->eta_bins_per_unit, c->phi_bins, c->eta_max);
    return to_status(hipGetLastError());
}

// the optional actors' states are complete for their bits
int check_actors(const dtr_actor_states* a) {
    if (!a) return DTR_SUCCESS;
    constexpr uint32_t kAll = DTR_ACTOR_MOMENTUM_ABORTER | DTR_ACTOR_TARGET_ABORTER | DTR_ACTOR_STEP_TRACER;
    if (a->actors & ~kAll) return DTR_ERROR_INVALID_ARGUMENT;
    if ((a->actors & DTR_ACTOR_MOMENTUM_ABORTER) && !(a->min_p == a->min_p && a->min_pT == a->min_pT))
        return DTR_ERROR_INVALID_ARGUMENT;
    if ((a->actors & DTR_ACTOR_TARGET_ABORTER) && !a->target_surface) return DTR_ERROR_INVALID_ARGUMENT;
    if ((a->actors & DTR_ACTOR_STEP_TRACER) && !a->trace.n) return DTR_ERROR_INVALID_ARGUMENT;
    return DTR_SUCCESS;
}

int check_propagate_args(const dtr_detector* det, const dtr_field* field, const dtr_prop_config* c,
                         const dtr_track_batch* in, const dtr_result_batch* out) {
    if (!det || !field || !c || !in || !out) return DTR_ERROR_INVALID_ARGUMENT;
    if (in->n == 0u) return DTR_SUCCESS;
    if (!in->free_params && !in->bound_params) return DTR_ERROR_INVALID_ARGUMENT;
    if (in->bound_params && !in->start_surface) return DTR_ERROR_INVALID_ARGUMENT;
    if (c->max_surfaces != 0u && (!out->surfaces || !out->n_surfaces)) return DTR_ERROR_INVALID_ARGUMENT;
    // the field-gradient Jacobian has no energy-loss terms
    if (det->volume_material && c->use_field_gradient) return DTR_ERROR_UNSUPPORTED;
    if (c->max_rk_updates == 0u) return DTR_ERROR_INVALID_ARGUMENT;
    return check_actors(c->actors);
}

int check_adjoint_args(const dtr_detector* det, const dtr_field* field, const dtr_prop_config* c,
                       const dtr_adjoint_config* a, const dtr_adjoint_batch* b, const dtr_hits* hits) {
    if (!det || !field || !c || !a || !b) return DTR_ERROR_INVALID_ARGUMENT;
    // the adjoint's steps follow vacuum volumes (no continuous energy loss
    // in the Runge-Kutta stages)
    if (det->volume_material) return DTR_ERROR_UNSUPPORTED;
    if (a->max_steps == 0u || a->max_events == 0u || a->checkpoint_every == 0u) return DTR_ERROR_INVALID_ARGUMENT;
    if (a->max_steps > (1u << 20) || a->max_events > 4096u || a->checkpoint_every > a->max_steps)
        return DTR_ERROR_INVALID_ARGUMENT;
    if (a->precision != DTR_PRECISION_FP64 && a->precision != DTR_PRECISION_MIXED) return DTR_ERROR_INVALID_ARGUMENT;
    if (a->model != DTR_ADJOINT_EXACT && a->model != DTR_ADJOINT_TRANSPORT) return DTR_ERROR_INVALID_ARGUMENT;
    if (a->block_size != 0u && (a->block_size % 64u != 0u || a->block_size > 256u)) return DTR_ERROR_INVALID_ARGUMENT;
    if (a->waves_per_simd > 2u) return DTR_ERROR_INVALID_ARGUMENT;
    if (!(a->sigma_loc0 > 0.0) || !(a->sigma_loc1 > 0.0)) return DTR_ERROR_INVALID_ARGUMENT;
    if (c->use_field_gradient) return DTR_ERROR_UNSUPPORTED;
    // the recording pass has its own actor chain (the tape replaces the
    // sequencer and the tracer)
    if (c->actors && c->actors->actors) return DTR_ERROR_UNSUPPORTED;
    if (c->max_rk_updates == 0u) return DTR_ERROR_INVALID_ARGUMENT;
    if (hits && (hits->max_hits == 0u || !hits->surface || !hits->loc || !hits->n_hits))
        return DTR_ERROR_INVALID_ARGUMENT;
    if (b->n == 0u) return DTR_SUCCESS;
    if (!b->bound_params || !b->start_surface || !b->tape) return DTR_ERROR_INVALID_ARGUMENT;
    if (b->tape_bytes < dtr::tape_bytes(b->n, a->max_steps, a->max_events, a->checkpoint_every))
        return DTR_ERROR_INVALID_ARGUMENT;
    return DTR_SUCCESS;
}

dtr::AdjointIO make_adjoint_io(const dtr_adjoint_config* a, const dtr_adjoint_batch* b, const dtr_hits* hits) {
    dtr::AdjointIO io{};
    io.bound0 = b->bound_params;
    io.start_sf = b->start_surface;
    io.n = b->n;
    if (hits) io.hits = dtr::HitsView{hits->max_hits, hits->surface, hits->loc, hits->n_hits};
    io.iv0 = 1.0 / (a->sigma_loc0 * a->sigma_loc0);
    io.iv1 = 1.0 / (a->sigma_loc1 * a->sigma_loc1);
    io.tape = dtr::tape_carve(b->tape, b->n, a->max_steps, a->max_events, a->checkpoint_every);
    return io;
}

// the adjoint's passes at one or two resident waves per SIMD
template <class P>
int adjoint_record_w(bool w2, const dtr_detector* det, const dtr_field* field, const dtr::DevConfig& dc,
                     const dtr::AdjointIO& io, uint32_t block, const uint32_t* order, hipStream_t st) {
    return w2 ? dtr::launch_adjoint_record<P, 2>(det, field, dc, io, block, order, st)
              : dtr::launch_adjoint_record<P, 1>(det, field, dc, io, block, order, st);
}
template <class P>
int adjoint_backward_w(bool w2, const dtr_detector* det, const dtr_field* field, const dtr::DevConfig& dc,
                       const dtr::AdjointIO& io, uint32_t block, const uint32_t* order, int exact, double* grad,
                       hipStream_t st) {
    return w2 ? dtr::launch_adjoint_backward<P, 2>(det, field, dc, io, block, order, exact, grad, st)
              : dtr::launch_adjoint_backward<P, 1>(det, field, dc, io, block, order, exact, grad, st);
}

// register budget: 1 or 2 resident waves per SIMD
template <class P>
int launch_fast(const dtr_detector* det, const dtr_field* field, const dtr::DevConfig& dc,
                const dtr_track_batch* in, const dtr::DevResults& r, const dtr_launch_config& L,
                hipStream_t stream) {
    if (L.waves_per_simd == 2u) return dtr::launch_batched<P, 2>(det, field, dc, in, r, L, stream);
    return dtr::launch_batched<P, 1>(det, field, dc, in, r, L, stream);
}

}  // namespace

namespace dtr {
size_t lds_bytes(uint32_t block) {
    return static_cast<size_t>(block) * kCache * 16u + (block / 64u) * sizeof(CoopShared);
}
}  // namespace dtr

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
    // every index the kernels follow is checked on the host first, so that
    // inconsistent tables are an error code, never an out-of-bounds read
    if (!dtr_host::validate_desc(*d).empty()) return DTR_ERROR_INVALID_ARGUMENT;
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
    if (!rc) rc = upload(d->material_maps, d->n_material_maps, &h->blocks[12]);
    if (!rc) {
        for (uint32_t k = 0; k < d->n_surfaces; ++k) {
            const dtr_surface& sf = d->surfaces[k];
            if (sf.shape >= DTR_LINE_SQUARE) h->wires = true;
            // a ring the parameter updater stops at needs the polar Jacobians
            if (sf.shape == DTR_RING2D && (sf.sf_id == DTR_SENSITIVE || sf.material != 0xffffffffu))
                h->wires = true;
        }
        for (uint32_t k = 0; k < d->n_materials; ++k) {
            if (d->materials[k].kind == DTR_MATERIAL_ROD) h->wires = true;
            if (d->materials[k].has_density) h->density_effect = true;
        }
        // the stepper's energy loss in volume material runs on the extended
        // geometry's kernels
        for (uint32_t k = 0; k < d->n_volumes; ++k)
            if (d->volumes[k].material != 0u) h->volume_material = h->wires = true;
        for (uint32_t k = 0; k < d->n_grids; ++k)
            if (d->grids[k].capacity > 1u) h->wires = true;
        if (!rc) {
            std::vector<dtr::NavSurface> nav(d->n_surfaces);
            for (uint32_t k = 0; k < d->n_surfaces; ++k) nav[k] = dtr::make_nav_surface(*d, k);
            // the cooperative search packs a bin index into 26 bits
            if (d->n_bins >= (1u << 26)) rc = DTR_ERROR_UNSUPPORTED;
            std::vector<dtr::NavRecord> bin_rec(d->n_bins), bf_rec(d->n_bf);
            std::vector<float> gtolk(d->n_grids);
            if (!rc) dtr::make_nav_records(*d, nav.data(), bin_rec.data(), bf_rec.data(), gtolk.data());
            if (!rc) rc = upload(nav.data(), d->n_surfaces, &h->blocks[8]);
            if (!rc) rc = upload(bin_rec.data(), d->n_bins, &h->blocks[9]);
            if (!rc) rc = upload(bf_rec.data(), d->n_bf, &h->blocks[10]);
            if (!rc) rc = upload(gtolk.data(), d->n_grids, &h->blocks[11]);
        }
    }
    if (!rc) {
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) rc = to_status(e);
        h->device = dev;
        h->n_cu = cus > 0 ? static_cast<uint32_t>(cus) : 1u;
    }
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
    h->dev.mat_maps = static_cast<const dtr_material_map*>(h->blocks[12]);
    h->fast.sf = static_cast<const dtr::NavSurface*>(h->blocks[8]);
    h->fast.masks = h->dev.masks;
    h->fast.materials = h->dev.materials;
    h->fast.mat_maps = h->dev.mat_maps;
    h->fast.volumes = h->dev.volumes;
    h->fast.grids = h->dev.grids;
    h->fast.transforms = h->dev.transforms;
    h->fast.bf = h->dev.bf;
    h->fast.bins = h->dev.bins;
    h->fast.bin_rec = static_cast<const dtr::NavRecord*>(h->blocks[9]);
    h->fast.bf_rec = static_cast<const dtr::NavRecord*>(h->blocks[10]);
    h->fast.grid_tolk = static_cast<const float*>(h->blocks[11]);
    h->n_volumes = d->n_volumes;
    h->n_surfaces = d->n_surfaces;
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
    const int rc = check_propagate_args(det, field, c, in, out);
    if (rc != DTR_SUCCESS || in->n == 0u) return rc;
    const dtr::DevConfig dc = dtr::make_dev_config(c);
    dtr::DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                      c->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                      out->n_trials, out->status};
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (c->use_field_gradient) {
        // the Jacobian with position columns runs on the batched engine's
        // FP64 policy (same results contract as the plain kernel)
        if (det->wires)
            return dtr::launch_lane<dtr::WithFieldGradient<dtr::WithExtendedGeometry<dtr::PolicyFP64>>>(
                det, field, dc, in, r, 64u, nullptr, 0, st);
        return dtr::launch_lane<dtr::WithFieldGradient<dtr::PolicyFP64>>(det, field, dc, in, r, 64u, nullptr, 0, st);
    }
    if (det->wires) return dtr::launch_plain<true>(det, field, dc, in, r, st);
    return dtr::launch_plain<false>(det, field, dc, in, r, st);
}

void dtr_default_actors(dtr_actor_states* a) {
    if (!a) return;
    std::memset(a, 0, sizeof(*a));
    // momentum_aborter::state (aborters.hpp:99-100): 10 MeV
    a->min_p = 10.f * 1e-3f;
    a->min_pT = 10.f * 1e-3f;
}

void dtr_default_launch(dtr_launch_config* l) {
    if (!l) return;
    std::memset(l, 0, sizeof(*l));
    l->precision = DTR_PRECISION_MIXED;
    l->schedule = DTR_SCHEDULE_PERSISTENT;
    l->block_size = 256u;
    l->waves_per_simd = 2u;
    l->tracks_per_claim = 64u;
    l->min_idle = 1u;
    l->navigation = DTR_NAVIGATION_WAVE;
    l->xcd_chunk = 64u;
}

void dtr_default_sort(dtr_sort_config* c) {
    if (!c) return;
    c->eta_bins_per_unit = 20u;
    c->phi_bins = 128u;
    c->eta_max = 5.f;
    c->pad = 0u;
}

size_t dtr_sort_workspace_size(const dtr_sort_config* c) {
    if (!c || c->eta_bins_per_unit == 0u || c->phi_bins == 0u || !(c->eta_max > 0.f)) return 0u;
    return sort_words(sort_cells(c)) * sizeof(uint32_t);
}

int dtr_sort_tracks(const dtr_sort_config* c, const double* fp, uint32_t n, uint32_t* order, void* workspace,
                    size_t workspace_bytes, void* stream) {
    if (!sort_config_ok(c)) return DTR_ERROR_INVALID_ARGUMENT;
    if (n == 0u) return DTR_SUCCESS;
    if (!fp || !order || !workspace) return DTR_ERROR_INVALID_ARGUMENT;
    const uint32_t cells = sort_cells(c);
    if (cells == 0u) return DTR_ERROR_INVALID_ARGUMENT;
    if (workspace_bytes < sort_words(cells) * sizeof(uint32_t)) return DTR_ERROR_INVALID_ARGUMENT;
    return sort_launch(c, fp, n, order, static_cast<uint32_t*>(workspace), cells, static_cast<hipStream_t>(stream));
}

int dtr_sort_tracks_static(const dtr_sort_config* c, const double* fp, uint32_t n, uint32_t* order, void* stream) {
    if (!sort_config_ok(c)) return DTR_ERROR_INVALID_ARGUMENT;
    const uint32_t cells = sort_cells(c);
    if (cells == 0u || cells > kStaticSortCells) return DTR_ERROR_INVALID_ARGUMENT;
    if (n == 0u) return DTR_SUCCESS;
    if (!fp || !order) return DTR_ERROR_INVALID_ARGUMENT;
    uint32_t* ws = nullptr;
    const int rc = static_scratch(&ws);
    if (rc != DTR_SUCCESS) return rc;
    return sort_launch(c, fp, n, order, ws, cells, static_cast<hipStream_t>(stream));
}

int dtr_propagate_ex(const dtr_detector* det, const dtr_field* field, const dtr_prop_config* c,
                     const dtr_track_batch* in, dtr_result_batch* out, const dtr_launch_config* launch,
                     void* stream) {
    const int rc = check_propagate_args(det, field, c, in, out);
    if (rc != DTR_SUCCESS) return rc;
    if (!launch) return DTR_ERROR_INVALID_ARGUMENT;
    dtr_launch_config L = *launch;
    if (L.block_size == 0u) L.block_size = 256u;
    if (L.waves_per_simd == 0u) L.waves_per_simd = 2u;
    if (L.tracks_per_claim == 0u) L.tracks_per_claim = 64u;
    if (L.min_idle == 0u) L.min_idle = 1u;
    if (L.min_idle > 64u || L.xcd_chunk > 65536u) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.block_size % 64u != 0u || L.block_size > 256u) return DTR_ERROR_INVALID_ARGUMENT;
    // the kernels are built for one or two resident waves per SIMD (512 or
    // 256 registers per lane); a persistent grid is sized with the same value
    if (L.tracks_per_claim % 64u != 0u || L.waves_per_simd > 2u) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.precision != DTR_PRECISION_FP64 && L.precision != DTR_PRECISION_MIXED) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.schedule != DTR_SCHEDULE_GRID && L.schedule != DTR_SCHEDULE_PERSISTENT) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.schedule == DTR_SCHEDULE_PERSISTENT && !L.queue) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.navigation > DTR_NAVIGATION_GATHER) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.schedule == DTR_SCHEDULE_PERSISTENT && L.navigation != DTR_NAVIGATION_WAVE) return DTR_ERROR_UNSUPPORTED;
    if (in->n == 0u) return DTR_SUCCESS;
    const dtr::DevConfig dc = dtr::make_dev_config(c);
    dtr::DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                      c->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                      out->n_trials, out->status};
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int gather = L.navigation == DTR_NAVIGATION_GATHER ? 1 : 0;
    if (det->volume_material) {
        // the energy loss in volume material runs on the per-lane kernels
        // (lane or gathered navigation)
        using VF = dtr::WithVolumeMaterial<dtr::WithExtendedGeometry<dtr::PolicyFP64>>;
        using VM = dtr::WithVolumeMaterial<dtr::WithExtendedGeometry<dtr::PolicyMixed>>;
        if (L.schedule != DTR_SCHEDULE_GRID || L.navigation == DTR_NAVIGATION_WAVE) return DTR_ERROR_UNSUPPORTED;
        if (L.precision == DTR_PRECISION_FP64)
            return dtr::launch_lane<VF>(det, field, dc, in, r, L.block_size, L.order, gather, st);
        return dtr::launch_lane<VM>(det, field, dc, in, r, L.block_size, L.order, gather, st);
    }
    if (det->wires) {
        // the extended geometry is compiled for two waves per SIMD
        if (c->use_field_gradient) {
            if (L.schedule != DTR_SCHEDULE_GRID || L.navigation == DTR_NAVIGATION_WAVE) return DTR_ERROR_UNSUPPORTED;
            if (L.precision == DTR_PRECISION_FP64)
                return dtr::launch_lane<dtr::WithFieldGradient<dtr::WithExtendedGeometry<dtr::PolicyFP64>>>(
                    det, field, dc, in, r, L.block_size, L.order, gather, st);
            return dtr::launch_lane<dtr::WithFieldGradient<dtr::WithExtendedGeometry<dtr::PolicyMixed>>>(
                det, field, dc, in, r, L.block_size, L.order, gather, st);
        }
        if (L.precision == DTR_PRECISION_FP64)
            return dtr::launch_batched<dtr::WithExtendedGeometry<dtr::PolicyFP64>, 2>(det, field, dc, in, r, L, st);
        return dtr::launch_batched<dtr::WithExtendedGeometry<dtr::PolicyMixed>, 2>(det, field, dc, in, r, L, st);
    }
    if (c->use_field_gradient) {
        if (L.schedule != DTR_SCHEDULE_GRID || L.navigation == DTR_NAVIGATION_WAVE) return DTR_ERROR_UNSUPPORTED;
        if (L.precision == DTR_PRECISION_FP64)
            return dtr::launch_lane<dtr::WithFieldGradient<dtr::PolicyFP64>>(det, field, dc, in, r, L.block_size,
                                                                             L.order, gather, st);
        return dtr::launch_lane<dtr::WithFieldGradient<dtr::PolicyMixed>>(det, field, dc, in, r, L.block_size,
                                                                          L.order, gather, st);
    }
    if (L.precision == DTR_PRECISION_FP64) return launch_fast<dtr::PolicyFP64>(det, field, dc, in, r, L, st);
    return launch_fast<dtr::PolicyMixed>(det, field, dc, in, r, L, st);
}

int dtr_propagate_direct(const dtr_detector* det, const dtr_field* field, const dtr_prop_config* c,
                         const dtr_track_batch* in, const dtr_surface_sequence* seq, dtr_result_batch* out,
                         const dtr_launch_config* launch, void* stream) {
    const int rc = check_propagate_args(det, field, c, in, out);
    if (rc != DTR_SUCCESS) return rc;
    if (!seq || !launch) return DTR_ERROR_INVALID_ARGUMENT;
    if (in->n == 0u) return DTR_SUCCESS;
    if (seq->max_surfaces == 0u || !seq->surfaces || !seq->n_surfaces) return DTR_ERROR_INVALID_ARGUMENT;
    if (seq->direction > DTR_DIRECTION_BACKWARD) return DTR_ERROR_INVALID_ARGUMENT;
    dtr_launch_config L = *launch;
    if (L.block_size == 0u) L.block_size = 64u;
    if (L.block_size % 64u != 0u || L.block_size > 256u) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.precision != DTR_PRECISION_FP64 && L.precision != DTR_PRECISION_MIXED) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.schedule != DTR_SCHEDULE_GRID || c->use_field_gradient) return DTR_ERROR_UNSUPPORTED;
    const dtr::DevConfig dc = dtr::make_dev_config(c);
    const dtr::DevConfig dd = dtr::direct_config(dc);
    dtr::DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                      c->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                      out->n_trials, out->status};
    const dtr::SeqView sq{seq->surfaces, seq->n_surfaces, seq->max_surfaces, det->n_surfaces,
                          seq->direction == DTR_DIRECTION_BACKWARD ? 1u : 0u};
    const uint32_t block = L.block_size;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (det->volume_material) {
        using VF = dtr::WithVolumeMaterial<dtr::WithExtendedGeometry<dtr::PolicyFP64>>;
        using VM = dtr::WithVolumeMaterial<dtr::WithExtendedGeometry<dtr::PolicyMixed>>;
        if (L.precision == DTR_PRECISION_FP64)
            return dtr::launch_direct<VF>(det, field, dc, dd, in, r, sq, block, L.order, st);
        return dtr::launch_direct<VM>(det, field, dc, dd, in, r, sq, block, L.order, st);
    }
    if (det->wires && L.precision == DTR_PRECISION_FP64)
        return dtr::launch_direct<dtr::WithExtendedGeometry<dtr::PolicyFP64>>(det, field, dc, dd, in, r, sq, block,
                                                                              L.order, st);
    if (det->wires)
        return dtr::launch_direct<dtr::WithExtendedGeometry<dtr::PolicyMixed>>(det, field, dc, dd, in, r, sq, block,
                                                                               L.order, st);
    if (L.precision == DTR_PRECISION_FP64)
        return dtr::launch_direct<dtr::PolicyFP64>(det, field, dc, dd, in, r, sq, block, L.order, st);
    return dtr::launch_direct<dtr::PolicyMixed>(det, field, dc, dd, in, r, sq, block, L.order, st);
}

size_t dtr_propagation_state_size(uint32_t n) {
    return static_cast<size_t>(n) * (dtr::kPauseD * sizeof(double) + dtr::kPauseW * sizeof(uint32_t));
}

int dtr_propagate_resumable(const dtr_detector* det, const dtr_field* field, const dtr_prop_config* c,
                            const dtr_track_batch* in, dtr_result_batch* out, const dtr_launch_config* launch,
                            const dtr_propagation_state* state, void* stream) {
    const int rc = check_propagate_args(det, field, c, in, out);
    if (rc != DTR_SUCCESS) return rc;
    if (!launch || !state) return DTR_ERROR_INVALID_ARGUMENT;
    dtr_launch_config L = *launch;
    if (L.block_size == 0u) L.block_size = 256u;
    if (L.block_size % 64u != 0u || L.block_size > 256u || L.xcd_chunk > 65536u) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.waves_per_simd > 2u) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.precision != DTR_PRECISION_FP64 && L.precision != DTR_PRECISION_MIXED) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.navigation > DTR_NAVIGATION_GATHER || L.schedule > DTR_SCHEDULE_PERSISTENT)
        return DTR_ERROR_INVALID_ARGUMENT;
    if (state->max_steps == 0u || state->resume > 1u) return DTR_ERROR_INVALID_ARGUMENT;
    if (L.schedule != DTR_SCHEDULE_GRID || L.navigation == DTR_NAVIGATION_WAVE) return DTR_ERROR_UNSUPPORTED;
    if (c->use_field_gradient) return DTR_ERROR_UNSUPPORTED;
    if (in->n == 0u) return DTR_SUCCESS;
    if (!state->data) return DTR_ERROR_INVALID_ARGUMENT;
    const dtr::DevConfig dc = dtr::make_dev_config(c);
    dtr::DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                      c->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                      out->n_trials, out->status};
    dtr::PauseView pv;
    pv.n = in->n;
    pv.d = static_cast<double*>(state->data);
    pv.w = reinterpret_cast<uint32_t*>(pv.d + static_cast<uint64_t>(dtr::kPauseD) * in->n);
    pv.max_steps = state->max_steps;
    pv.resume = state->resume;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int gather = L.navigation == DTR_NAVIGATION_GATHER ? 1 : 0;
    const bool fp64 = L.precision == DTR_PRECISION_FP64;
    using namespace dtr;
    if (det->volume_material) {
        if (fp64)
            return launch_resumable<WithVolumeMaterial<WithExtendedGeometry<PolicyFP64>>>(
                det, field, dc, in, r, L.block_size, L.order, gather, L.xcd_chunk, pv, st);
        return launch_resumable<WithVolumeMaterial<WithExtendedGeometry<PolicyMixed>>>(
            det, field, dc, in, r, L.block_size, L.order, gather, L.xcd_chunk, pv, st);
    }
    if (det->wires) {
        if (fp64)
            return launch_resumable<WithExtendedGeometry<PolicyFP64>>(det, field, dc, in, r, L.block_size, L.order,
                                                                      gather, L.xcd_chunk, pv, st);
        return launch_resumable<WithExtendedGeometry<PolicyMixed>>(det, field, dc, in, r, L.block_size, L.order,
                                                                   gather, L.xcd_chunk, pv, st);
    }
    if (fp64)
        return launch_resumable<PolicyFP64>(det, field, dc, in, r, L.block_size, L.order, gather, L.xcd_chunk, pv,
                                            st);
    return launch_resumable<PolicyMixed>(det, field, dc, in, r, L.block_size, L.order, gather, L.xcd_chunk, pv,
                                         st);
}

void dtr_default_adjoint(dtr_adjoint_config* a) {
    if (!a) return;
    std::memset(a, 0, sizeof(*a));
    a->max_steps = 512u;
    a->max_events = 64u;
    a->checkpoint_every = 32u;
    a->precision = DTR_PRECISION_MIXED;
    a->model = DTR_ADJOINT_EXACT;
    a->block_size = 64u;
    a->waves_per_simd = 2u;
    a->sigma_loc0 = 0.010;
    a->sigma_loc1 = 0.050;
}

size_t dtr_adjoint_tape_size(const dtr_adjoint_config* a, uint32_t n) {
    if (!a || a->max_steps == 0u || a->max_events == 0u || a->checkpoint_every == 0u) return 0u;
    return dtr::tape_bytes(n, a->max_steps, a->max_events, a->checkpoint_every);
}

int dtr_adjoint_forward(const dtr_detector* det, const dtr_field* field, const dtr_prop_config* c,
                        const dtr_adjoint_config* a, const dtr_adjoint_batch* b, const dtr_hits* hits,
                        dtr_hits* record, double* chi2, double* grad_transport, uint32_t* status, void* stream) {
    const int rc = check_adjoint_args(det, field, c, a, b, hits);
    if (rc != DTR_SUCCESS || b->n == 0u) return rc;
    if (record && (record->max_hits == 0u || !record->surface || !record->loc || !record->n_hits))
        return DTR_ERROR_INVALID_ARGUMENT;
    dtr::DevConfig dc = dtr::make_dev_config(c);
    dc.max_surfaces = 0u;  // the event list on the tape replaces the surface sequencer
    // chi2 and its gradient do not depend on the covariance: it is only
    // carried when the navigation reads it (noise estimation) or the
    // Jacobian chain is asked for
    if (!grad_transport && !c->estimate_scattering_noise) dc.do_cov = 0u;
    dtr::AdjointIO io = make_adjoint_io(a, b, hits);
    if (record) io.record = dtr::HitsOut{record->max_hits, record->surface, record->loc, record->n_hits};
    io.chi2 = chi2;
    io.grad_transport = grad_transport;
    io.status = status;
    const uint32_t block = a->block_size ? a->block_size : 128u;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const bool w2 = a->waves_per_simd != 1u;
    if (det->wires) {
        if (a->precision == DTR_PRECISION_FP64)
            return adjoint_record_w<dtr::WithExtendedGeometry<dtr::PolicyFP64>>(w2, det, field, dc, io, block,
                                                                                b->order, st);
        return adjoint_record_w<dtr::WithExtendedGeometry<dtr::PolicyMixed>>(w2, det, field, dc, io, block, b->order,
                                                                             st);
    }
    if (a->precision == DTR_PRECISION_FP64)
        return adjoint_record_w<dtr::PolicyFP64>(w2, det, field, dc, io, block, b->order, st);
    return adjoint_record_w<dtr::PolicyMixed>(w2, det, field, dc, io, block, b->order, st);
}

int dtr_adjoint_backward(const dtr_detector* det, const dtr_field* field, const dtr_prop_config* c,
                         const dtr_adjoint_config* a, const dtr_adjoint_batch* b, const dtr_hits* hits,
                         double* grad, void* stream) {
    const int rc = check_adjoint_args(det, field, c, a, b, hits);
    if (rc != DTR_SUCCESS || b->n == 0u) return rc;
    if (!grad) return DTR_ERROR_INVALID_ARGUMENT;
    dtr::DevConfig dc = dtr::make_dev_config(c);
    dc.max_surfaces = 0u;
    const dtr::AdjointIO io = make_adjoint_io(a, b, hits);
    const uint32_t block = a->block_size ? a->block_size : 128u;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int exact = a->model == DTR_ADJOINT_EXACT ? 1 : 0;
    const bool w2 = a->waves_per_simd != 1u;
    if (det->wires) {
        if (a->precision == DTR_PRECISION_FP64)
            return adjoint_backward_w<dtr::WithExtendedGeometry<dtr::PolicyFP64>>(w2, det, field, dc, io, block,
                                                                                  b->order, exact, grad, st);
        return adjoint_backward_w<dtr::WithExtendedGeometry<dtr::PolicyMixed>>(w2, det, field, dc, io, block,
                                                                               b->order, exact, grad, st);
    }
    if (a->precision == DTR_PRECISION_FP64)
        return adjoint_backward_w<dtr::PolicyFP64>(w2, det, field, dc, io, block, b->order, exact, grad, st);
    return adjoint_backward_w<dtr::PolicyMixed>(w2, det, field, dc, io, block, b->order, exact, grad, st);
}

}  // extern "C"
