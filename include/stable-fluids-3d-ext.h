/*
 * stable-fluids-3d-ext.h -- additive extensions of the stable-fluids C ABI.
 *
 * None of these has a counterpart in the reference interface (stable-fluids-3d.h); the
 * base header and its seven entry points are unchanged. They expose what the base ABI
 * cannot express: setting the flow state (velocity seeding, checkpoint/resume), choosing
 * the smoother, reading residual norms and timing single stages.
 */
#ifndef STABLE_FLUIDS_3D_EXT_H
#define STABLE_FLUIDS_3D_EXT_H

#include "stable-fluids-3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum StableFluidsExtSmoother {
    STABLE_FLUIDS_EXT_SMOOTHER_RBGS   = 0, /* red-black Gauss-Seidel, the reference smoother (default) */
    STABLE_FLUIDS_EXT_SMOOTHER_JACOBI = 1, /* Jacobi, every cell from the previous iterate */
} StableFluidsExtSmoother;

/* Storage of the smoother iterate. F32 is the reference path. The other values run
 * every diffusion and pressure solve of K passes as ceil(K/m) refinement cycles: an FP32
 * defect pass r = rhs - A x, then up to m smoother passes on A e = r from e = 0 with the
 * correction e stored in the named format (scaled per cycle by a power of two >= max|r|),
 * then x += e. In exact arithmetic the cycles equal the K plain passes. F32_REFINED runs
 * the same cycles with an FP32 correction and exists to check them. Not available with
 * the RBGS smoother on a periodic axis of odd extent. */
typedef enum StableFluidsExtIterateStorage {
    STABLE_FLUIDS_EXT_STORAGE_F32         = 0,
    STABLE_FLUIDS_EXT_STORAGE_F16         = 1,
    STABLE_FLUIDS_EXT_STORAGE_BF16        = 2,
    STABLE_FLUIDS_EXT_STORAGE_F32_REFINED = 3,
} StableFluidsExtIterateStorage;

typedef enum StableFluidsExtWarmStart {
    STABLE_FLUIDS_EXT_WARM_START_ZERO    = 0, /* pressure restarts from zero each step (reference) */
    /* The first pressure iterate is the warm-start network's prediction (see
     * stable-fluids-unet.h) from the half-resolution divergence, followed by
     * warm_start_iterations smoother passes. When the guess residual ||b - A p0|| exceeds
     * 10x the zero-guess residual ||b||, the guess is dropped and the step runs the plain
     * pressure_iterations passes from zero (bit-identical to the reference path).
     * Needs extents multiples of 2^(c+2) (c = warm_start_coarsening), FP32 iterate storage and a model loaded with
     * stable_fluids_ext_set_pressure_model_cuda (until then the network is all zeros). */
    STABLE_FLUIDS_EXT_WARM_START_LEARNED = 1,
} StableFluidsExtWarmStart;


typedef struct StableFluidsExtOptions {
    uint32_t struct_size;                /* sizeof(StableFluidsExtOptions) */
    uint32_t smoother;                   /* StableFluidsExtSmoother */
    uint32_t iterate_storage;            /* StableFluidsExtIterateStorage */
    uint32_t defect_correction_interval; /* m, passes per refinement cycle; 0 = 10, at most 4096 */
    uint32_t pressure_warm_start;        /* StableFluidsExtWarmStart */
    uint32_t warm_start_iterations;      /* passes after a learned guess (<= pressure_iterations) */
    uint32_t warm_start_coarsening;      /* the network runs on the grid coarsened 2^c times per
                                            axis (c = 1, 2 or 3; 0 = 1); extents multiples of 2^(c+2) */
    uint32_t flags;                      /* reserved, must be zero */
} StableFluidsExtOptions;

/* Same as stable_fluids_create_context_cuda, with options (NULL = reference behaviour).
 * Unsupported option values return STABLE_FLUIDS_RESULT_BACKEND_FAILURE. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_create_context_cuda(const StableFluidsContextCreateDesc* desc, const StableFluidsExtOptions* options, void** out_context, StableFluidsScalarFieldHandle* out_scalar_field_handles, uint32_t out_scalar_field_handle_capacity, StableFluidsVectorFieldHandle* out_vector_field_handles, uint32_t out_vector_field_handle_capacity);

/* Sets the flow velocity from three device arrays (NULL = zeros), asynchronously on the
 * context stream. Needed to start from a non-zero velocity and to resume a checkpoint. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_flow_velocity_cuda(void* context, const float* values_x, const float* values_y, const float* values_z);

/* Animated force computed on the device inside the step graph. The reference's plume scene
 * recomputes its swirl and drift forcing on the host and uploads force x and z before every
 * step (vulkan-app/scene_plume.cpp:323-335, masks built at :241-289); this replaces that
 * per-step upload. Before the force of step n is added (n counts from start_step, one per
 * step; the context has one counter, which every call sets to its start_step):
 *   phase = n * phase_rate                      (n converted to float)
 *   ax = drift_x_amplitude * sin(phase),  az = drift_z_amplitude * cos(phase * drift_z_rate)
 *   s  = swirl_amplitude * cos(phase * swirl_rate)
 *   force.x = ax * drift + s * swirl_x,  force.z = az * drift + s * swirl_z
 * in FP32, in this operation order, into components x and z of force field `field` (its y
 * component keeps what was set). The scene's constants are phase_rate 0.045, amplitudes
 * 1.15 / 0.85 / 1.65, rates 0.71 / 0.53. The masks are device arrays of nx*ny*nz floats,
 * copied by the call. NULL `animation` turns the animation off. Synchronous: it waits for
 * the context stream and rebuilds the step graph. Contexts with an animation cannot join a
 * batch, and a context that is in a batch must not be given one (the batch's own step graph
 * would not run it). */
typedef struct StableFluidsExtForceAnimation {
    float phase_rate;
    float drift_x_amplitude;
    float drift_z_amplitude;
    float drift_z_rate;
    float swirl_amplitude;
    float swirl_rate;
    uint32_t start_step;
} StableFluidsExtForceAnimation;

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_force_animation_cuda(void* context, StableFluidsVectorFieldHandle field, const StableFluidsExtForceAnimation* animation, const float* drift, const float* swirl_x, const float* swirl_z);

/* Step counter of the force animation (the n of the next step). Synchronises the context
 * stream. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_force_animation_step_cuda(void* context, uint32_t* out_step);

/* Sets the pressure field from a device array (NULL = zeros). The next step overwrites it
 * unless a warm start uses it; provided for checkpoint/resume of every viewable field. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_flow_pressure_cuda(void* context, const float* values);

typedef struct StableFluidsExtNorms {
    double divergence_l2;   /* sqrt(sum div^2) over all cells, FP64 accumulation */
    double divergence_max;  /* max |div| */
    double velocity_l2;     /* sqrt(sum |v|^2) */
    double pressure_l2;     /* sqrt(sum p^2) */
} StableFluidsExtNorms;

/* Reduces the current divergence, velocity and pressure views on the GPU (FP64
 * accumulation) and waits for the result. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_field_norms_cuda(void* context, StableFluidsExtNorms* out_norms);

/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * reduction of stable_fluids_ext_field_norms_cuda over caller-given device fields of
 * `cells` floats each (for example a context's views), bit-identical to it on the same
 * fields. It allocates nothing and does not synchronise: the partial sums live in a
 * static device array of the library (32,800 bytes, one per GPU, held until the process
 * ends), and the result is written to out_norms_device (device memory) on `stream`
 * (NULL: the null stream), which must belong to the current device. Every cell count
 * fits the array (at most 1024 partial sums); NULL pointers, cells = 0 and a stream of
 * another device return STABLE_FLUIDS_RESULT_BACKEND_FAILURE. All calls on one GPU share
 * its array: calls that may overlap on one GPU (unordered streams, threads) race and give
 * wrong results; order them on one stream. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_field_norms_noalloc_cuda(const float* divergence, const float* velocity_x, const float* velocity_y, const float* velocity_z, const float* pressure, uint64_t cells, void* stream, StableFluidsExtNorms* out_norms_device);

typedef enum StableFluidsExtStage {
    /* one smoother pass of the pressure solve as the step runs it: for a plain FP32 RBGS
     * solve a middle pass that reads b and the black cells and writes the black cells
     * (8 B/cell); otherwise the full-iterate pass */
    STABLE_FLUIDS_EXT_STAGE_PRESSURE_PASS  = 0,
    STABLE_FLUIDS_EXT_STAGE_VELOCITY_PASS  = 1, /* the same for one velocity component's diffusion */
    STABLE_FLUIDS_EXT_STAGE_ADVECT         = 2, /* velocity self-advection */
    STABLE_FLUIDS_EXT_STAGE_PROJECT        = 3, /* projection + divergence + |v| */
    STABLE_FLUIDS_EXT_STAGE_JACOBI_PASS    = 4, /* one Jacobi pass of the pressure solve, whatever the smoother option */
    STABLE_FLUIDS_EXT_STAGE_COPY           = 5, /* plain copy of one field (8 B/cell), to calibrate byte counters */
    /* 6, 7: not used */
    STABLE_FLUIDS_EXT_STAGE_FULL_PRESSURE_PASS = 8, /* one full-iterate pass (read x_old, b; write x_new: 12 B/cell) of the pressure solve */
    STABLE_FLUIDS_EXT_STAGE_FULL_VELOCITY_PASS = 9, /* the same for one velocity component's diffusion */
    /* whole sub-sequences of the step, each exactly as the step enqueues it (several
     * kernels: stable_fluids_ext_stage_bytes_cuda gives 0 for them) */
    STABLE_FLUIDS_EXT_STAGE_PRESSURE_SOLVE = 10, /* pressure solve from the last divergence (learned or plain) */
    STABLE_FLUIDS_EXT_STAGE_PROJECTION     = 11, /* divergence + pressure solve + projection of the diffused velocity */
} StableFluidsExtStage;

/* Enqueues `repeat` launches of one stage on the context stream, outside the step graph
 * and with the step's own buffers and shapes, for per-kernel (or per-sub-sequence)
 * timing. The simulation state is left undefined afterwards (re-seed before stepping
 * again), except after PRESSURE_SOLVE and PROJECTION, which recompute what the step's
 * own sub-sequence computes from the same inputs. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_run_stage_cuda(void* context, uint32_t stage, uint32_t repeat);

/* Algorithmic HBM bytes one launch of `stage` moves (each array read once and written
 * once), as used for roofline figures. */
STABLE_FLUIDS_API uint64_t stable_fluids_ext_stage_bytes_cuda(void* context, uint32_t stage);

/* Tile shapes (x cells by rows, workgroup threads) of the RBGS passes that carry only
 * black cells between passes (plain FP32 RBGS solves on even nx). */
typedef enum StableFluidsExtTile {
    STABLE_FLUIDS_EXT_TILE_128X16 = 0, /* 512 threads (default) */
    STABLE_FLUIDS_EXT_TILE_64X32  = 1, /* 512 threads */
    STABLE_FLUIDS_EXT_TILE_64X16  = 2, /* 256 threads */
    STABLE_FLUIDS_EXT_TILE_32X32  = 3, /* 256 threads */
    STABLE_FLUIDS_EXT_TILE_32X16  = 4, /* 128 threads */
} StableFluidsExtTile;

/* Sets the tile shape and z-chunk length (planes per workgroup, 0 = the default plan) of
 * the context's RBGS passes and rebuilds the step graph (synchronous). Tiling never
 * changes results. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_smoother_tiling_cuda(void* context, uint32_t tile, uint32_t z_chunk);
/* The tile shape and the z-chunk length in use (either pointer may be NULL). */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_get_smoother_tiling_cuda(void* context, uint32_t* out_tile, uint32_t* out_z_chunk);
/* Tiling autotuner for one context, in place of the reference's occupancy-derived block
 * size (backend-cuda.cu:490-515): times `repeats` RBGS pressure passes (writing only
 * scratch buffers) for every tile shape with z chunks {default, 4, 6, 8, 12, 16, 32} and keeps
 * the fastest. A shape found in cache_path is taken from it, and the file is rewritten
 * with every entry (cache_path may be NULL; the format is shared with the batch
 * autotuner, single contexts under members 0). Rebuilds the step graph (synchronous);
 * a context whose pressure solve does not run the black-cell passes is left as it is.
 * out_tile / out_z_chunk (either may be NULL) = the tiling in use afterwards. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_autotune_cuda(void* context, const char* cache_path, uint32_t repeats, uint32_t* out_tile, uint32_t* out_z_chunk);

/* Returns the context's HIP stream (hipStream_t as void*). */
STABLE_FLUIDS_API void* stable_fluids_ext_get_stream_cuda(void* context);

/* ---- batched simulations ----
 * A batch steps many contexts (uneven grid sizes allowed) with one kernel launch per
 * stage instead of one per stage per context: one graph launch runs every context's step
 * in the reference step order, each context's result bit-identical to
 * stable_fluids_step_cuda on it. The contexts must agree on diffuse_iterations and
 * pressure_iterations (both > 0), scalar and vector field counts and the smoother, have
 * positive viscosity and scalar diffusivities, FP32 storage, no learned warm start, and
 * (RBGS) even extents on periodic axes. The batch reads the contexts' buffers in place:
 * views, setters and updates keep working between batch steps; do not step a context
 * on its own stream while a batch step runs. stream NULL: the batch owns a new stream. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_create_cuda(void* const* contexts, uint32_t count, void* stream, void** out_batch);
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_destroy_cuda(void* batch);
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_step_cuda(void* batch);
STABLE_FLUIDS_API void* stable_fluids_ext_batch_get_stream_cuda(void* batch);
/* Smoother z-chunk length of one context in the batch (0 = default plan); rebuilds the
 * batch graph. Tiling never changes results. A batch starts from each context's own
 * smoother tiling. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_set_z_chunk_cuda(void* batch, uint32_t sim, uint32_t z_chunk);
/* Tile shape (StableFluidsExtTile; used by the black-cell passes of even-nx contexts) and
 * z-chunk length of one context in the batch; rebuilds the batch graph. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_set_tiling_cuda(void* batch, uint32_t sim, uint32_t tile, uint32_t z_chunk);
/* out_tiling4 = {tiles_x, tiles_y, z chunks, z chunk length} of context `sim`'s pressure
 * passes; out_launches_per_step = kernel launches in one batch step. Either may be NULL. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_info_cuda(void* batch, uint32_t sim, int32_t* out_tiling4, uint32_t* out_launches_per_step);
/* Tile shape of context `sim` in the batch. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_get_tile_cuda(void* batch, uint32_t sim, uint32_t* out_tile);
/* Autotuner, in place of the reference's occupancy-derived block size
 * (backend-cuda.cu:490-515). For every grid shape in the batch it times `repeats`
 * rounds of batched smoother passes in the step's mix (one three-component velocity
 * pass, three single-field passes; writing only scratch buffers) over that shape's
 * contexts alone for each candidate tiling -- every tile shape with z chunks {default, 4, 6,
 * 8, 12, 16, 32} for black-cell passes, z chunks {default, 2, 4, ..., 64} otherwise -- taking
 * the minimum of three interleaved rounds. It then times whole-batch passes under: each
 * shape on its fastest tiling; all shapes on one tile shape (one launch per pass), each
 * on its fastest chunk for it, for every tile shape; the starting tiling; keeps the
 * fastest, then re-chooses each shape's z chunk (tile kept) by whole-batch passes.
 * Shapes listed in cache_path (a line "stable-fluids-tiling 2", then lines "nx
 * ny nz smoother members tile zc") are taken from it, and the file is rewritten with all
 * entries (cache_path may be NULL). Rebuilds the batch graph; *out_tuned_shapes =
 * shapes timed (not from the cache). */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_batch_autotune_cuda(void* batch, const char* cache_path, uint32_t repeats, uint32_t* out_tuned_shapes);

/* ---- hand-off to a Vulkan renderer (Linux opaque file descriptors) ----
 * The compute side of the reference viewer's capture slots (vulkan-app/app.cpp:816-871,
 * the fd branch; snapshot export vulkan-app/app.ixx:240-273 and scene_plume.cpp:343-360).
 * import: takes a VkDeviceMemory exported as VK_EXTERNAL_MEMORY_HANDLE_TYPE_OPAQUE_FD_BIT
 * (memory_bytes = its allocation size) and maps buffer_bytes of it from buffer_offset;
 * semaphore_fd (or -1 for none) is a timeline VkSemaphore exported as
 * VK_EXTERNAL_SEMAPHORE_HANDLE_TYPE_OPAQUE_FD_BIT. On success the runtime owns both
 * descriptors (the caller must not close them). On failure the slot is released; a
 * descriptor imported before the failing step (memory, then semaphore) is consumed.
 * export_view: copies the view `request` names into the slot's buffer (SOA3 views as x,
 * then y, then z) on request->consumer_stream (NULL: the context stream), then signals
 * the semaphore to signal_value on that stream. Asynchronous: the renderer waits on the
 * semaphore value (the reference also synchronizes the stream after signalling).
 * Fails if the view does not fit the buffer. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_interop_import_cuda(int memory_fd, uint64_t memory_bytes, uint64_t buffer_offset, uint64_t buffer_bytes, int semaphore_fd, void** out_slot);
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_interop_destroy_cuda(void* slot);
/* the slot's mapped device buffer (NULL for a NULL slot) */
STABLE_FLUIDS_API void* stable_fluids_ext_interop_buffer_cuda(void* slot);
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_interop_export_view_cuda(void* context, void* slot, const StableFluidsViewRequest* request, uint64_t signal_value);

/* Device array [nz][ny][nx] FP32 of the last step's pre-projection divergence (the
 * pressure right-hand side; the DIVERGENCE view holds the post-projection field). Valid
 * until the context is destroyed; NULL for a NULL context. */
STABLE_FLUIDS_API const float* stable_fluids_ext_get_pressure_rhs_cuda(void* context);

/* Loads the warm-start network parameters (host FP32, stable-fluids-unet.h layout, count =
 * stable_fluids_unet_param_count()) into a context created with
 * STABLE_FLUIDS_EXT_WARM_START_LEARNED. Synchronous; takes effect from the next step. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_set_pressure_model_cuda(void* context, const float* params, uint64_t count);

/* Warm-start outcome of the last step: out[0] = ||b - A p0|| (guess), out[1] = ||b||
 * (zero guess), out[2] = 1 when the step fell back to the plain passes. Synchronises the
 * context stream. The fallback passes skipped by an accepted guess run as one launch whose
 * workgroups wait only for finished neighbour blocks, so it needs no exclusive use of the
 * GPU; should one of its waits ever run past its bound (about 0.4 s), the step's pressure
 * is filled with NaN and this call returns BACKEND_FAILURE. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_warm_start_stats_cuda(void* context, double* out3);

/* Training pair of the warm-start network from the last step's pressure system: the
 * pre-projection divergence div and the exact solution p* of A p = -h^2 div (pressure
 * boundary rules; FP64 conjugate gradients to ||r|| <= rel_tol ||b||, at most
 * max_iterations, started from the previous call's solution; zero-mean when the system
 * has no outflow face), both averaged over blocks of f^3 cells, f = 2^coarsening
 * (extents multiples of f). Writes input[v][2] = (div_c / sigma, wall mask) and
 * target[v] = p*_c / ((f h)^2 sigma), sigma = RMS(div_c), on the coarse grid, device
 * FP32. Synchronous. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_training_pair_cuda(void* context, float* input, float* target, uint32_t coarsening, double rel_tol, uint32_t max_iterations, double* out_rel_residual, uint32_t* out_iterations);

#ifdef __cplusplus
}
#endif

#endif /* STABLE_FLUIDS_3D_EXT_H */
