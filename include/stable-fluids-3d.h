/*
 * stable-fluids-3d.h -- public C ABI of the MI355X (gfx950) stable-fluids 3D step.
 *
 * Drop-in boundary: every type, enum value, struct layout and entry-point name below
 * matches the reference interface in stable-fluids-3d.h:22-154 so that existing callers
 * (benchmark.cu, vulkan-app/scene_plume.cpp) relink against libstable-fluids.so unchanged.
 * The "_cuda" suffix is kept because it is part of the exported symbol names those callers
 * bind; the implementation is HIP on gfx950 only. Additive features live in
 * stable-fluids-3d-ext.h.
 *
 * Conventions (reference backend-cuda.cu:481-800):
 *  - every field is a packed float array of nx*ny*nz values, x fastest:
 *    index = (z*ny + y)*nx + x; cell (x,y,z) has its centre at ((x+.5)h, (y+.5)h, (z+.5)h);
 *  - pointers passed to update_* are DEVICE pointers, copied asynchronously on the
 *    context stream; NULL means "fill with zeros";
 *  - get_view returns library-owned device pointers (no copy); contents change at the
 *    next step;
 *  - all work is asynchronous on the context stream (the caller's hipStream_t passed as
 *    void*, or a non-blocking stream owned by the library when NULL);
 *  - results: 0 ok, 1 out of memory, 2 any other failure. No error strings.
 */
#ifndef STABLE_FLUIDS_3D_H
#define STABLE_FLUIDS_3D_H

#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define STABLE_FLUIDS_API __attribute__((visibility("default")))
#else
#define STABLE_FLUIDS_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* reference stable-fluids-3d.h:22-26 */
typedef enum StableFluidsResult {
    STABLE_FLUIDS_RESULT_OK              = 0,
    STABLE_FLUIDS_RESULT_OUT_OF_MEMORY   = 1,
    STABLE_FLUIDS_RESULT_BACKEND_FAILURE = 2,
} StableFluidsResult;

/* reference stable-fluids-3d.h:28-33. PERIODIC must be set on both faces of an axis. */
typedef enum StableFluidsFlowBoundaryType {
    STABLE_FLUIDS_FLOW_BOUNDARY_NO_SLIP_WALL   = 0,
    STABLE_FLUIDS_FLOW_BOUNDARY_FREE_SLIP_WALL = 1,
    STABLE_FLUIDS_FLOW_BOUNDARY_OUTFLOW        = 2,
    STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC       = 3,
} StableFluidsFlowBoundaryType;

/* reference stable-fluids-3d.h:35-39 */
typedef enum StableFluidsScalarBoundaryType {
    STABLE_FLUIDS_SCALAR_BOUNDARY_FIXED_VALUE = 0,
    STABLE_FLUIDS_SCALAR_BOUNDARY_ZERO_FLUX   = 1,
    STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC    = 2,
} StableFluidsScalarBoundaryType;

/* 1-based handles; 0 is never valid (reference backend-cuda.cu:544,551,695). */
typedef uint32_t StableFluidsScalarFieldHandle;
typedef uint32_t StableFluidsVectorFieldHandle;

/* reference stable-fluids-3d.h:44-50: wall velocity (NO_SLIP / FREE_SLIP normal
 * component) and the Dirichlet pressure used by OUTFLOW faces. */
typedef struct StableFluidsFlowBoundaryFaceDesc {
    uint32_t type;
    float velocity_x;
    float velocity_y;
    float velocity_z;
    float pressure;
} StableFluidsFlowBoundaryFaceDesc;

/* reference stable-fluids-3d.h:52-59 */
typedef struct StableFluidsFlowBoundaryConfig {
    StableFluidsFlowBoundaryFaceDesc x_minus;
    StableFluidsFlowBoundaryFaceDesc x_plus;
    StableFluidsFlowBoundaryFaceDesc y_minus;
    StableFluidsFlowBoundaryFaceDesc y_plus;
    StableFluidsFlowBoundaryFaceDesc z_minus;
    StableFluidsFlowBoundaryFaceDesc z_plus;
} StableFluidsFlowBoundaryConfig;

/* reference stable-fluids-3d.h:61-64; value is used by FIXED_VALUE faces. */
typedef struct StableFluidsScalarBoundaryFaceDesc {
    uint32_t type;
    float value;
} StableFluidsScalarBoundaryFaceDesc;

/* reference stable-fluids-3d.h:66-73 */
typedef struct StableFluidsScalarBoundaryConfig {
    StableFluidsScalarBoundaryFaceDesc x_minus;
    StableFluidsScalarBoundaryFaceDesc x_plus;
    StableFluidsScalarBoundaryFaceDesc y_minus;
    StableFluidsScalarBoundaryFaceDesc y_plus;
    StableFluidsScalarBoundaryFaceDesc z_minus;
    StableFluidsScalarBoundaryFaceDesc z_plus;
} StableFluidsScalarBoundaryConfig;

/* reference stable-fluids-3d.h:75-85. Frozen at create time. */
typedef struct StableFluidsSimulationConfig {
    int32_t nx;
    int32_t ny;
    int32_t nz;
    float cell_size;
    float dt;
    float viscosity;
    int32_t diffuse_iterations;
    int32_t pressure_iterations;
    StableFluidsFlowBoundaryConfig flow_boundary;
} StableFluidsSimulationConfig;

/* reference stable-fluids-3d.h:86-91 */
typedef struct StableFluidsScalarFieldDesc {
    float diffusion;
    float dissipation;
    float initial_value;
    StableFluidsScalarBoundaryConfig boundary;
} StableFluidsScalarFieldDesc;

/* reference stable-fluids-3d.h:93-95 */
typedef enum StableFluidsVectorFieldUsage {
    STABLE_FLUIDS_VECTOR_FIELD_FORCE = 0,
} StableFluidsVectorFieldUsage;

/* reference stable-fluids-3d.h:97-102 */
typedef struct StableFluidsVectorFieldDesc {
    StableFluidsVectorFieldUsage usage;
    float initial_value_x;
    float initial_value_y;
    float initial_value_z;
} StableFluidsVectorFieldDesc;

/* reference stable-fluids-3d.h:104-111; stream is a hipStream_t or NULL. */
typedef struct StableFluidsContextCreateDesc {
    StableFluidsSimulationConfig config;
    void* stream;
    const StableFluidsScalarFieldDesc* scalar_fields;
    uint32_t scalar_field_count;
    const StableFluidsVectorFieldDesc* vector_fields;
    uint32_t vector_field_count;
} StableFluidsContextCreateDesc;

/* replaces reference stable-fluids-3d.h:112 (implementation backend-cuda.cu:481-680).
 * Handles are written as index+1 for index < capacity. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_create_context_cuda(const StableFluidsContextCreateDesc* desc, void** out_context, StableFluidsScalarFieldHandle* out_scalar_field_handles, uint32_t out_scalar_field_handle_capacity, StableFluidsVectorFieldHandle* out_vector_field_handles, uint32_t out_vector_field_handle_capacity);
/* replaces reference stable-fluids-3d.h:113 (backend-cuda.cu:682-690); synchronises the stream first. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_destroy_context_cuda(void* context);

/* replaces reference stable-fluids-3d.h:115 (backend-cuda.cu:692-699) */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_update_scalar_field_cuda(void* context, StableFluidsScalarFieldHandle field, const float* values);
/* replaces reference stable-fluids-3d.h:116 (backend-cuda.cu:701-708) */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_update_scalar_field_source_cuda(void* context, StableFluidsScalarFieldHandle field, const float* values);
/* replaces reference stable-fluids-3d.h:117 (backend-cuda.cu:710-719) */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_update_vector_field_cuda(void* context, StableFluidsVectorFieldHandle field, const float* values_x, const float* values_y, const float* values_z);
/* replaces reference stable-fluids-3d.h:118 (backend-cuda.cu:721-726): one hipGraphLaunch. */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_step_cuda(void* context);

/* reference stable-fluids-3d.h:120-128 */
typedef enum StableFluidsViewKind {
    STABLE_FLUIDS_VIEW_SCALAR_FIELD_DATA       = 0,
    STABLE_FLUIDS_VIEW_SCALAR_FIELD_SOURCE     = 1,
    STABLE_FLUIDS_VIEW_VECTOR_FIELD            = 2,
    STABLE_FLUIDS_VIEW_FLOW_VELOCITY           = 3,
    STABLE_FLUIDS_VIEW_FLOW_VELOCITY_MAGNITUDE = 4,
    STABLE_FLUIDS_VIEW_FLOW_PRESSURE           = 5,
    STABLE_FLUIDS_VIEW_FLOW_DIVERGENCE         = 6,
} StableFluidsViewKind;

/* reference stable-fluids-3d.h:130-133 */
typedef enum StableFluidsViewLayout {
    STABLE_FLUIDS_VIEW_LAYOUT_F32_3D      = 0,
    STABLE_FLUIDS_VIEW_LAYOUT_F32_3D_SOA3 = 1,
} StableFluidsViewLayout;

/* reference stable-fluids-3d.h:135-140; consumer_stream is a hipStream_t or NULL. */
typedef struct StableFluidsViewRequest {
    uint32_t kind;
    StableFluidsScalarFieldHandle scalar_field;
    StableFluidsVectorFieldHandle vector_field;
    void* consumer_stream;
} StableFluidsViewRequest;

/* reference stable-fluids-3d.h:142-152 */
typedef struct StableFluidsView {
    uint32_t layout;
    int32_t nx;
    int32_t ny;
    int32_t nz;
    uint64_t row_stride_bytes;
    uint64_t slice_stride_bytes;
    const float* data0;
    const float* data1;
    const float* data2;
} StableFluidsView;

/* replaces reference stable-fluids-3d.h:154 (backend-cuda.cu:728-800) */
STABLE_FLUIDS_API StableFluidsResult stable_fluids_get_view_cuda(void* context, const StableFluidsViewRequest* request, StableFluidsView* out_view);

#ifdef __cplusplus
}
#endif

#endif /* STABLE_FLUIDS_3D_H */
