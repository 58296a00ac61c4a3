// stable-fluids-benchmark -- the reference's benchmark driver contract (benchmark.cu:15-265)
// run against libstable-fluids.so: same command-line flags (:29-59), same scene (:72-155:
// 128x160x128, periodic x/z, no-slip floor, outflow top, ellipsoidal plume with a swirling
// force, one density field with diffusion and dissipation) and the same two output lines
// (:237-254), so its mlups figure is directly comparable. Only the C ABI of
// include/stable-fluids-3d.h is used; device buffers, the stream and events are HIP.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/stable-fluids-3d.h"

namespace {

struct Options {
    int nx = 128, ny = 160, nz = 128;
    int warmup = 48, steps = 256;
    int diffuse_iterations = 24, pressure_iterations = 96;
    float cell_size = 0.01f, dt = 1.0f / 90.0f, viscosity = 0.00012f;
    float field_diffusion = 0.00005f, field_dissipation = 0.35f;
};

bool parse(int argc, char** argv, Options& o, bool& help) {
    help = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--help") {
            help = true;
            return true;
        }
        if (i + 1 >= argc) {
            std::fprintf(stderr, "missing value for %s\n", a.c_str());
            return false;
        }
        const char* v = argv[++i];
        if (a == "--nx") o.nx = std::atoi(v);
        else if (a == "--ny") o.ny = std::atoi(v);
        else if (a == "--nz") o.nz = std::atoi(v);
        else if (a == "--warmup") o.warmup = std::atoi(v);
        else if (a == "--steps") o.steps = std::atoi(v);
        else if (a == "--diffuse-iters") o.diffuse_iterations = std::atoi(v);
        else if (a == "--pressure-iters") o.pressure_iterations = std::atoi(v);
        else if (a == "--cell-size") o.cell_size = std::strtof(v, nullptr);
        else if (a == "--dt") o.dt = std::strtof(v, nullptr);
        else if (a == "--viscosity") o.viscosity = std::strtof(v, nullptr);
        else if (a == "--field-diffusion") o.field_diffusion = std::strtof(v, nullptr);
        else if (a == "--field-dissipation") o.field_dissipation = std::strtof(v, nullptr);
        else {
            std::fprintf(stderr, "unknown argument: %s\n", a.c_str());
            return false;
        }
    }
    return true;
}

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "%s failed: %s\n", what, hipGetErrorString(e));
    return false;
}

bool sf_ok(StableFluidsResult r, const char* what) {
    if (r == STABLE_FLUIDS_RESULT_OK) return true;
    std::fprintf(stderr, "%s failed: %d\n", what, static_cast<int>(r));
    return false;
}

StableFluidsFlowBoundaryFaceDesc flow_face(uint32_t type) { return StableFluidsFlowBoundaryFaceDesc{type, 0.0f, 0.0f, 0.0f, 0.0f}; }
StableFluidsScalarBoundaryFaceDesc scalar_face(uint32_t type) { return StableFluidsScalarBoundaryFaceDesc{type, 0.0f}; }

} // namespace

int main(int argc, char** argv) {
    Options o;
    bool help = false;
    if (!parse(argc, argv, o, help)) return EXIT_FAILURE;
    if (help) {
        std::printf("stable-fluids-benchmark [--nx N] [--ny N] [--nz N] [--warmup N] [--steps N]\n"
                    "                        [--diffuse-iters N] [--pressure-iters N]\n"
                    "                        [--cell-size H] [--dt DT] [--viscosity V]\n"
                    "                        [--field-diffusion D] [--field-dissipation K]\n");
        return EXIT_SUCCESS;
    }

    // scene: ellipsoidal plume near the floor (benchmark.cu:72-155)
    const size_t cells = static_cast<size_t>(o.nx) * o.ny * o.nz;
    const float ex = static_cast<float>(o.nx) * o.cell_size, ey = static_cast<float>(o.ny) * o.cell_size, ez = static_cast<float>(o.nz) * o.cell_size;
    const float cx = ex * 0.50f, cy = ey * 0.16f, cz = ez * 0.50f;
    const float rx = ex * 0.07f, ry = ey * 0.06f, rz = ez * 0.07f;
    std::vector<float> fx(cells, 0.0f), fy(cells, 0.0f), fz(cells, 0.0f), src(cells, 0.0f);
    for (int z = 0; z < o.nz; ++z)
        for (int y = 0; y < o.ny; ++y)
            for (int x = 0; x < o.nx; ++x) {
                const float px = (static_cast<float>(x) + 0.5f) * o.cell_size;
                const float py = (static_cast<float>(y) + 0.5f) * o.cell_size;
                const float pz = (static_cast<float>(z) + 0.5f) * o.cell_size;
                const float dx = (px - cx) / rx, dy = (py - cy) / ry, dz = (pz - cz) / rz;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > 1.0f) continue;
                const size_t i    = (static_cast<size_t>(z) * o.ny + y) * o.nx + x;
                const float plume = std::exp(-2.2f * r2);
                src[i]            = 28.0f * plume;
                fx[i]             = 1.8f * plume * std::sin((pz / (ez + 1.0e-6f)) * 6.2831853f);
                fy[i]             = 7.5f * plume;
                fz[i]             = 1.4f * plume * std::cos((px / (ex + 1.0e-6f)) * 6.2831853f);
            }

    StableFluidsSimulationConfig cfg{};
    cfg.nx = o.nx;
    cfg.ny = o.ny;
    cfg.nz = o.nz;
    cfg.cell_size           = o.cell_size;
    cfg.dt                  = o.dt;
    cfg.viscosity           = o.viscosity;
    cfg.diffuse_iterations  = o.diffuse_iterations;
    cfg.pressure_iterations = o.pressure_iterations;
    cfg.flow_boundary       = {flow_face(STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC), flow_face(STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC),
                               flow_face(STABLE_FLUIDS_FLOW_BOUNDARY_NO_SLIP_WALL), flow_face(STABLE_FLUIDS_FLOW_BOUNDARY_OUTFLOW),
                               flow_face(STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC), flow_face(STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC)};
    StableFluidsScalarFieldDesc density{};
    density.diffusion     = o.field_diffusion;
    density.dissipation   = o.field_dissipation;
    density.initial_value = 0.0f;
    density.boundary      = {scalar_face(STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC), scalar_face(STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC),
                             scalar_face(STABLE_FLUIDS_SCALAR_BOUNDARY_ZERO_FLUX), scalar_face(STABLE_FLUIDS_SCALAR_BOUNDARY_ZERO_FLUX),
                             scalar_face(STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC), scalar_face(STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC)};
    const StableFluidsVectorFieldDesc force{STABLE_FLUIDS_VECTOR_FIELD_FORCE, 0.0f, 0.0f, 0.0f};

    hipStream_t stream = nullptr;
    if (!hip_ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags")) return EXIT_FAILURE;
    StableFluidsContextCreateDesc desc{};
    desc.config             = cfg;
    desc.stream             = stream;
    desc.scalar_fields      = &density;
    desc.scalar_field_count = 1;
    desc.vector_fields      = &force;
    desc.vector_field_count = 1;
    void* ctx               = nullptr;
    StableFluidsScalarFieldHandle sh = 0;
    StableFluidsVectorFieldHandle vh = 0;
    if (!sf_ok(stable_fluids_create_context_cuda(&desc, &ctx, &sh, 1, &vh, 1), "stable_fluids_create_context_cuda")) return EXIT_FAILURE;

    const size_t bytes = cells * sizeof(float);
    float *dfx = nullptr, *dfy = nullptr, *dfz = nullptr, *dsrc = nullptr;
    for (float** p : {&dfx, &dfy, &dfz, &dsrc})
        if (!hip_ok(hipMalloc(reinterpret_cast<void**>(p), bytes), "hipMalloc")) return EXIT_FAILURE;
    if (!hip_ok(hipMemcpyAsync(dfx, fx.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") || !hip_ok(hipMemcpyAsync(dfy, fy.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") ||
        !hip_ok(hipMemcpyAsync(dfz, fz.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") || !hip_ok(hipMemcpyAsync(dsrc, src.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync"))
        return EXIT_FAILURE;
    if (!sf_ok(stable_fluids_update_vector_field_cuda(ctx, vh, dfx, dfy, dfz), "stable_fluids_update_vector_field_cuda")) return EXIT_FAILURE;
    if (!sf_ok(stable_fluids_update_scalar_field_source_cuda(ctx, sh, dsrc), "stable_fluids_update_scalar_field_source_cuda")) return EXIT_FAILURE;

    // warmup, then the timed steps bracketed by events on the context stream (:197-215)
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (!hip_ok(hipEventCreate(&t0), "hipEventCreate") || !hip_ok(hipEventCreate(&t1), "hipEventCreate")) return EXIT_FAILURE;
    for (int i = 0; i < o.warmup; ++i)
        if (!sf_ok(stable_fluids_step_cuda(ctx), "stable_fluids_step_cuda warmup")) return EXIT_FAILURE;
    if (!hip_ok(hipStreamSynchronize(stream), "hipStreamSynchronize warmup")) return EXIT_FAILURE;
    if (!hip_ok(hipEventRecord(t0, stream), "hipEventRecord")) return EXIT_FAILURE;
    for (int i = 0; i < o.steps; ++i)
        if (!sf_ok(stable_fluids_step_cuda(ctx), "stable_fluids_step_cuda")) return EXIT_FAILURE;
    if (!hip_ok(hipEventRecord(t1, stream), "hipEventRecord") || !hip_ok(hipEventSynchronize(t1), "hipEventSynchronize")) return EXIT_FAILURE;
    float total_ms = 0.0f;
    if (!hip_ok(hipEventElapsedTime(&total_ms, t0, t1), "hipEventElapsedTime")) return EXIT_FAILURE;

    // export the density through a view (:217-230)
    std::vector<float> rho(cells, 0.0f);
    StableFluidsView view{};
    const StableFluidsViewRequest req{STABLE_FLUIDS_VIEW_SCALAR_FIELD_DATA, sh, 0, stream};
    if (!sf_ok(stable_fluids_get_view_cuda(ctx, &req, &view), "stable_fluids_get_view_cuda")) return EXIT_FAILURE;
    if (!hip_ok(hipMemcpyAsync(rho.data(), view.data0, bytes, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync density") || !hip_ok(hipStreamSynchronize(stream), "hipStreamSynchronize export")) return EXIT_FAILURE;

    const double avg_step_ms = o.steps > 0 ? static_cast<double>(total_ms) / o.steps : 0.0;
    const double mlups       = total_ms > 0.0f ? static_cast<double>(cells) * o.steps / (static_cast<double>(total_ms) * 1000.0) : 0.0;
    const float total        = std::accumulate(rho.begin(), rho.end(), 0.0f);
    const float peak         = rho.empty() ? 0.0f : *std::max_element(rho.begin(), rho.end());
    std::printf("benchmark grid=%dx%dx%d warmup=%d steps=%d dt=%.6f viscosity=%.7f field_diffusion=%.7f field_dissipation=%.4f\n", o.nx, o.ny, o.nz, o.warmup, o.steps, o.dt, o.viscosity, o.field_diffusion, o.field_dissipation);
    std::printf("timing total_ms=%.3f avg_step_ms=%.6f mlups=%.3f total_density=%.6f peak_density=%.6f\n", static_cast<double>(total_ms), avg_step_ms, mlups, total, peak);

    (void)hipEventDestroy(t1);
    (void)hipEventDestroy(t0);
    for (float* p : {dsrc, dfz, dfy, dfx}) (void)hipFree(p);
    stable_fluids_destroy_context_cuda(ctx);
    (void)hipStreamDestroy(stream);
    return EXIT_SUCCESS;
}
