// sf_interop.hip -- hands field views to a Vulkan renderer through external memory and a
// timeline semaphore (stable-fluids-3d-ext.h). The consumer side of the reference's
// viewer: it imports each capture slot's device memory and timeline semaphore into the
// compute API (vulkan-app/app.cpp:816-871, the opaque-fd branch) and, per snapshot, copies
// a view into the slot and signals the semaphore on the simulation's stream
// (vulkan-app/app.ixx:240-273, scene_plume.cpp:343-360). Host code only: the copy is a
// device-to-device memcpy and the signal a stream operation.
#include <hip/hip_runtime.h>

#include <new>

#include "../../include/stable-fluids-3d-ext.h"
#include "../../include/stable-fluids-3d.h"

namespace {

struct InteropSlot {
    hipExternalMemory_t memory       = nullptr;
    hipExternalSemaphore_t semaphore = nullptr;
    void* buffer                     = nullptr; // mapped [offset, offset + bytes) of the memory
    uint64_t bytes                   = 0;

    ~InteropSlot() {
        // a mapped buffer is freed like an allocation, before its memory object goes
        if (buffer) (void)hipFree(buffer);
        if (semaphore) (void)hipDestroyExternalSemaphore(semaphore);
        if (memory) (void)hipDestroyExternalMemory(memory);
    }
};

} // namespace

extern "C" {

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_interop_import_cuda(int memory_fd, uint64_t memory_bytes, uint64_t buffer_offset, uint64_t buffer_bytes, int semaphore_fd, void** out_slot) {
    if (out_slot == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    *out_slot = nullptr;
    if (memory_fd < 0 || memory_bytes == 0 || buffer_bytes == 0 || buffer_offset > memory_bytes || buffer_bytes > memory_bytes - buffer_offset) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    InteropSlot* slot = new (std::nothrow) InteropSlot{};
    if (slot == nullptr) return STABLE_FLUIDS_RESULT_OUT_OF_MEMORY;
    hipExternalMemoryHandleDesc md{};
    md.type      = hipExternalMemoryHandleTypeOpaqueFd;
    md.handle.fd = memory_fd;
    md.size      = memory_bytes;
    hipError_t e = hipImportExternalMemory(&slot->memory, &md);
    if (e == hipSuccess) {
        hipExternalMemoryBufferDesc bd{};
        bd.offset = buffer_offset;
        bd.size   = buffer_bytes;
        e         = hipExternalMemoryGetMappedBuffer(&slot->buffer, slot->memory, &bd);
    }
    if (e == hipSuccess && semaphore_fd >= 0) {
        hipExternalSemaphoreHandleDesc sd{};
        sd.type      = hipExternalSemaphoreHandleTypeTimelineSemaphoreFd;
        sd.handle.fd = semaphore_fd;
        e            = hipImportExternalSemaphore(&slot->semaphore, &sd);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        delete slot;
        return e == hipErrorOutOfMemory ? STABLE_FLUIDS_RESULT_OUT_OF_MEMORY : STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }
    slot->bytes = buffer_bytes;
    *out_slot   = slot;
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_interop_destroy_cuda(void* slot) {
    if (slot == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    delete static_cast<InteropSlot*>(slot);
    return STABLE_FLUIDS_RESULT_OK;
}

STABLE_FLUIDS_API void* stable_fluids_ext_interop_buffer_cuda(void* slot) { return slot == nullptr ? nullptr : static_cast<InteropSlot*>(slot)->buffer; }

STABLE_FLUIDS_API StableFluidsResult stable_fluids_ext_interop_export_view_cuda(void* context, void* slot, const StableFluidsViewRequest* request, uint64_t signal_value) {
    if (context == nullptr || slot == nullptr || request == nullptr) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    auto& s = *static_cast<InteropSlot*>(slot);
    StableFluidsView view{};
    const StableFluidsResult r = stable_fluids_get_view_cuda(context, request, &view);
    if (r != STABLE_FLUIDS_RESULT_OK) return r;
    const uint64_t field  = static_cast<uint64_t>(view.nx) * static_cast<uint64_t>(view.ny) * static_cast<uint64_t>(view.nz) * sizeof(float);
    const int parts       = view.layout == STABLE_FLUIDS_VIEW_LAYOUT_F32_3D_SOA3 ? 3 : 1;
    if (field * parts > s.bytes) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    // the copy and the signal run on the consumer stream, which the view is handed off to
    hipStream_t st        = static_cast<hipStream_t>(request->consumer_stream != nullptr ? request->consumer_stream : stable_fluids_ext_get_stream_cuda(context));
    const void* src[3]    = {view.data0, view.data1, view.data2};
    for (int i = 0; i < parts; ++i)
        if (hipMemcpyAsync(static_cast<char*>(s.buffer) + field * static_cast<uint64_t>(i), src[i], field, hipMemcpyDeviceToDevice, st) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    if (s.semaphore != nullptr) {
        hipExternalSemaphoreSignalParams p{};
        p.params.fence.value = signal_value;
        if (hipSignalExternalSemaphoresAsync(&s.semaphore, &p, 1, st) != hipSuccess) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
    }
    return STABLE_FLUIDS_RESULT_OK;
}

} // extern "C"
