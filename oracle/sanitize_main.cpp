// sanitize_main.cpp -- drives the host-loop oracle (sf_oracle.cpp) through its C API for a
// build with AddressSanitizer and UndefinedBehaviorSanitizer (host code only):
//
//   g++ -std=c++17 -O1 -g -fsanitize=address,undefined -fno-sanitize-recover=all \
//       -fno-omit-frame-pointer -ffp-contract=off sanitize_main.cpp sf_oracle.cpp -o sanitize
//
// TEST INFRASTRUCTURE ONLY (tests/test_oracle_sanitize.py builds and runs it). Every
// boundary type on every face, odd and one-cell extents, both precisions and smoothers,
// dissipation, zero iteration counts, a large-CFL field that traces far outside the grid,
// and the per-stage entry points. Exits 0 when every run finished with finite fields.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sf_oracle.h"

namespace {

using FlowFace   = StableFluidsFlowBoundaryFaceDesc;
using ScalarFace = StableFluidsScalarBoundaryFaceDesc;

StableFluidsFlowBoundaryConfig flow(const uint32_t types[6]) {
    StableFluidsFlowBoundaryConfig b{};
    FlowFace* f[6] = {&b.x_minus, &b.x_plus, &b.y_minus, &b.y_plus, &b.z_minus, &b.z_plus};
    for (int i = 0; i < 6; ++i) *f[i] = FlowFace{types[i], 0.1f * static_cast<float>(i), -0.05f, 0.2f, 0.25f * static_cast<float>(i & 1)};
    return b;
}

StableFluidsScalarBoundaryConfig scalar(const uint32_t types[6]) {
    StableFluidsScalarBoundaryConfig b{};
    ScalarFace* f[6] = {&b.x_minus, &b.x_plus, &b.y_minus, &b.y_plus, &b.z_minus, &b.z_plus};
    for (int i = 0; i < 6; ++i) *f[i] = ScalarFace{types[i], 0.5f - 0.2f * static_cast<float>(i)};
    return b;
}

// a deterministic field in [-scale, scale]
std::vector<float> field(const size_t n, const unsigned seed, const float scale) {
    std::vector<float> v(n);
    unsigned s = seed * 2654435761u + 1u;
    for (auto& x : v) {
        s = s * 1664525u + 1013904223u;
        x = scale * (static_cast<float>(s >> 8) / 8388608.0f - 1.0f);
    }
    return v;
}

bool finite(const std::vector<double>& v) {
    for (const double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

int run(const int nx, const int ny, const int nz, const uint32_t ft[6], const uint32_t st[6], const int precision, const int smoother, const int d_iters, const int p_iters, const float vel, const int steps) {
    StableFluidsScalarFieldDesc sd[2] = {};
    sd[0].diffusion     = 2e-3f;
    sd[0].dissipation   = 0.3f;
    sd[0].initial_value = 0.25f;
    sd[0].boundary      = scalar(st);
    sd[1].diffusion     = 0.0f;
    sd[1].boundary      = scalar(st);
    StableFluidsVectorFieldDesc vd[1] = {};
    vd[0].usage                       = STABLE_FLUIDS_VECTOR_FIELD_FORCE;
    vd[0].initial_value_y             = 0.5f;
    StableFluidsContextCreateDesc desc{};
    desc.config.nx                  = nx;
    desc.config.ny                  = ny;
    desc.config.nz                  = nz;
    desc.config.cell_size           = 1.0f / 16.0f;
    desc.config.dt                  = 0.05f;
    desc.config.viscosity           = 1e-3f;
    desc.config.diffuse_iterations  = d_iters;
    desc.config.pressure_iterations = p_iters;
    desc.config.flow_boundary       = flow(ft);
    desc.scalar_fields              = sd;
    desc.scalar_field_count         = 2;
    desc.vector_fields              = vd;
    desc.vector_field_count         = 1;
    void* sim                       = sfo_create(&desc, precision, smoother);
    if (sim == nullptr) {
        std::fprintf(stderr, "sfo_create failed for %dx%dx%d\n", nx, ny, nz);
        return 1;
    }
    const size_t n = static_cast<size_t>(nx) * ny * nz;
    const auto vx = field(n, 1, vel), vy = field(n, 2, vel), vz = field(n, 3, vel), src = field(n, 4, 1.0f);
    int bad = sfo_set_flow_velocity(sim, vx.data(), vy.data(), vz.data()) != 0;
    bad |= sfo_update_scalar_field_source(sim, 1, src.data()) != 0;
    bad |= sfo_update_vector_field(sim, 1, vz.data(), vx.data(), vy.data()) != 0;
    sfo_step(sim, steps);
    std::vector<double> a(n), b(n), c(n);
    for (const uint32_t kind : {STABLE_FLUIDS_VIEW_FLOW_VELOCITY, STABLE_FLUIDS_VIEW_FLOW_PRESSURE, STABLE_FLUIDS_VIEW_FLOW_DIVERGENCE, STABLE_FLUIDS_VIEW_FLOW_VELOCITY_MAGNITUDE}) {
        bad |= sfo_get_field(sim, kind, 0, a.data(), b.data(), c.data()) != 0;
        bad |= !finite(a);
    }
    for (uint32_t h = 1; h <= 2; ++h) {
        bad |= sfo_get_field(sim, STABLE_FLUIDS_VIEW_SCALAR_FIELD_DATA, h, a.data(), nullptr, nullptr) != 0;
        bad |= !finite(a);
    }
    sfo_destroy(sim);
    if (bad) std::fprintf(stderr, "run %dx%dx%d p%d s%d failed\n", nx, ny, nz, precision, smoother);
    return bad;
}

// the per-stage entry points on one grid, samples far outside it included
int stages(const int nx, const int ny, const int nz, const uint32_t ft[6], const uint32_t st[6]) {
    const auto fb = flow(ft);
    const auto sb = scalar(st);
    const size_t n = static_cast<size_t>(nx) * ny * nz;
    std::vector<double> f(n), g(n), h(n), o(n), out3(3);
    for (size_t i = 0; i < n; ++i) {
        f[i] = std::sin(0.3 * static_cast<double>(i));
        g[i] = std::cos(0.7 * static_cast<double>(i));
        h[i] = 0.5 * std::sin(1.1 * static_cast<double>(i));
    }
    double acc = 0.0;
    for (int z = -2; z <= nz + 1; ++z)
        for (int y = -2; y <= ny + 1; ++y)
            for (int x = -2; x <= nx + 1; ++x) {
                acc += sfo_load_scalar(f.data(), x, y, z, nx, ny, nz, &sb) + sfo_load_pressure(f.data(), x, y, z, nx, ny, nz, &fb);
                for (int c = 0; c < 3; ++c) acc += sfo_load_velocity(f.data(), c, x, y, z, nx, ny, nz, &fb);
            }
    for (const double q : {-3.7, -0.01, 0.3, 1.9, 40.0}) {
        acc += sfo_sample_scalar(f.data(), q, 0.5 * q, -q, nx, ny, nz, 0.1, &sb);
        for (int c = 0; c < 3; ++c) acc += sfo_sample_velocity(f.data(), c, -q, q, 0.25 * q, nx, ny, nz, 0.1, &fb);
        sfo_trace_rk2(f.data(), g.data(), h.data(), q, -q, 2.0 * q, 3.0, nx, ny, nz, 0.1, &fb, out3.data());
        acc += out3[0] + out3[1] + out3[2];
    }
    for (int c = 0; c < 3; ++c) sfo_advect_velocity(o.data(), f.data(), c, g.data(), h.data(), f.data(), 5.0, nx, ny, nz, 0.1, &fb);
    sfo_advect_scalar(o.data(), f.data(), g.data(), h.data(), f.data(), 5.0, nx, ny, nz, 0.1, &sb, &fb);
    for (const int sm : {SFO_SMOOTHER_RBGS, SFO_SMOOTHER_JACOBI}) {
        sfo_smooth_scalar(o.data(), f.data(), 0.2, 3, sm, nx, ny, nz, &sb);
        for (int c = 0; c < 3; ++c) sfo_smooth_velocity(o.data(), g.data(), 0.2, c, 3, sm, nx, ny, nz, &fb);
        sfo_smooth_pressure(o.data(), h.data(), 0.01, 3, sm, nx, ny, nz, &fb);
    }
    sfo_divergence(o.data(), f.data(), g.data(), h.data(), nx, ny, nz, 0.1, &fb);
    sfo_project(f.data(), g.data(), h.data(), o.data(), nx, ny, nz, 0.1, &fb);
    const bool ok = std::isfinite(acc) && finite(o) && finite(f);
    if (!ok) std::fprintf(stderr, "stages %dx%dx%d failed\n", nx, ny, nz);
    return ok ? 0 : 1;
}

} // namespace

int main() {
    constexpr uint32_t NS = STABLE_FLUIDS_FLOW_BOUNDARY_NO_SLIP_WALL, FS = STABLE_FLUIDS_FLOW_BOUNDARY_FREE_SLIP_WALL;
    constexpr uint32_t OF = STABLE_FLUIDS_FLOW_BOUNDARY_OUTFLOW, PF = STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC;
    constexpr uint32_t FV = STABLE_FLUIDS_SCALAR_BOUNDARY_FIXED_VALUE, ZF = STABLE_FLUIDS_SCALAR_BOUNDARY_ZERO_FLUX;
    constexpr uint32_t PS = STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC;
    const uint32_t walls_f[6] = {NS, FS, OF, NS, FS, OF}, walls_s[6] = {FV, ZF, FV, ZF, ZF, FV};
    const uint32_t per_f[6] = {PF, PF, PF, PF, PF, PF}, per_s[6] = {PS, PS, PS, PS, PS, PS};
    const uint32_t mix_f[6] = {PF, PF, NS, OF, PF, PF}, mix_s[6] = {PS, PS, FV, ZF, PS, PS};
    struct Grid {
        int nx, ny, nz;
    };
    const Grid grids[] = {{9, 7, 6}, {8, 4, 5}, {1, 6, 4}, {5, 1, 3}, {4, 3, 1}, {1, 1, 1}, {2, 2, 2}};
    int bad = 0;
    for (const auto& g : grids) {
        for (const int precision : {64, 32})
            for (const int smoother : {SFO_SMOOTHER_RBGS, SFO_SMOOTHER_JACOBI}) {
                bad |= run(g.nx, g.ny, g.nz, walls_f, walls_s, precision, smoother, 3, 4, 0.5f, 2);
                bad |= run(g.nx, g.ny, g.nz, per_f, per_s, precision, smoother, 2, 3, 0.5f, 2);
                bad |= run(g.nx, g.ny, g.nz, mix_f, mix_s, precision, smoother, 2, 3, 0.5f, 2);
            }
        bad |= run(g.nx, g.ny, g.nz, mix_f, mix_s, 64, SFO_SMOOTHER_RBGS, 0, 0, 0.5f, 2); // no passes at all
        bad |= run(g.nx, g.ny, g.nz, walls_f, walls_s, 64, SFO_SMOOTHER_RBGS, 2, 2, 300.0f, 2); // traces far outside
        bad |= stages(g.nx, g.ny, g.nz, walls_f, walls_s);
        bad |= stages(g.nx, g.ny, g.nz, per_f, per_s);
        bad |= stages(g.nx, g.ny, g.nz, mix_f, mix_s);
    }
    std::printf(bad ? "sanitize: FAILED\n" : "sanitize: all runs clean\n");
    return bad ? 1 : 0;
}
