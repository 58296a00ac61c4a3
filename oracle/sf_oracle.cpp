/*
 * sf_oracle.cpp -- host-loop oracle of the stable-fluids 3D step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libstable-fluids.so, never part of it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this.
 *
 * Every function restates one piece of the reference (file:line in each comment) as a
 * plain loop over cells, templated on the arithmetic type R (double = parity oracle,
 * float = same operations in single precision). OpenMP splits the z range; a red or black
 * half-sweep only reads cells of the other colour (or the cell itself through a wall
 * ghost), so the parallel loop gives the serial result.
 *
 * Half-sweep semantics: every cell of one colour is computed from the values present
 * before that half-sweep. This equals the reference whenever the reference is race free
 * (every axis non-periodic or of even length). On an odd-length periodic axis the
 * reference kernel races (cells 0 and n-1 share a colour and are neighbours through the
 * wrap); the snapshot rule is one of the orders that race can produce and is the one
 * the HIP kernels implement.
 *
 * Pinning: tests/test_oracle_*.py check each function against closed forms and
 * independent formulations (dense/sparse matrix solves, DCT solves, cyclic shifts,
 * hand-computed ghost tables); the reference ships no golden vectors.
 */
#include "sf_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <new>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace sfo {

enum : uint32_t {
    kNoSlip = STABLE_FLUIDS_FLOW_BOUNDARY_NO_SLIP_WALL,
    kFreeSlip = STABLE_FLUIDS_FLOW_BOUNDARY_FREE_SLIP_WALL,
    kOutflow = STABLE_FLUIDS_FLOW_BOUNDARY_OUTFLOW,
    kFlowPeriodic = STABLE_FLUIDS_FLOW_BOUNDARY_PERIODIC,
    kFixed = STABLE_FLUIDS_SCALAR_BOUNDARY_FIXED_VALUE,
    kZeroFlux = STABLE_FLUIDS_SCALAR_BOUNDARY_ZERO_FLUX,
    kScalarPeriodic = STABLE_FLUIDS_SCALAR_BOUNDARY_PERIODIC,
};

// reference backend-cuda.cu:58-60 (x fastest)
inline std::size_t at(int x, int y, int z, int nx, int ny) {
    return (static_cast<std::size_t>(z) * static_cast<std::size_t>(ny) + static_cast<std::size_t>(y)) * static_cast<std::size_t>(nx) + static_cast<std::size_t>(x);
}

inline const StableFluidsFlowBoundaryFaceDesc& face_of(const StableFluidsFlowBoundaryConfig& b, int axis, bool minus) {
    switch (axis) {
    case 0: return minus ? b.x_minus : b.x_plus;
    case 1: return minus ? b.y_minus : b.y_plus;
    default: return minus ? b.z_minus : b.z_plus;
    }
}
inline const StableFluidsScalarBoundaryFaceDesc& face_of(const StableFluidsScalarBoundaryConfig& b, int axis, bool minus) {
    switch (axis) {
    case 0: return minus ? b.x_minus : b.x_plus;
    case 1: return minus ? b.y_minus : b.y_plus;
    default: return minus ? b.z_minus : b.z_plus;
    }
}
inline bool periodic(const StableFluidsFlowBoundaryConfig& b, int a) {
    return face_of(b, a, true).type == kFlowPeriodic && face_of(b, a, false).type == kFlowPeriodic;
}
inline bool periodic(const StableFluidsScalarBoundaryConfig& b, int a) {
    return face_of(b, a, true).type == kScalarPeriodic && face_of(b, a, false).type == kScalarPeriodic;
}
inline int wrap(int c, int n) {
    c %= n;
    return c < 0 ? c + n : c;
}

// reference backend-cuda.cu:62-94: axes in x, y, z order; periodic wraps and continues,
// ZERO_FLUX clamps and continues, FIXED_VALUE returns the face value at once.
template <class R>
R load_scalar(const R* f, int x, int y, int z, int nx, int ny, int nz, const StableFluidsScalarBoundaryConfig& bc) {
    int c[3] = {x, y, z};
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < n[a]) continue;
        const auto& face = face_of(bc, a, c[a] < 0);
        if (periodic(bc, a) && n[a] > 0)
            c[a] = wrap(c[a], n[a]);
        else if (face.type == kZeroFlux && n[a] > 0)
            c[a] = c[a] < 0 ? 0 : n[a] - 1;
        else
            return static_cast<R>(face.value);
    }
    return f[at(c[0], c[1], c[2], nx, ny)];
}

// reference backend-cuda.cu:96-128: periodic wraps; OUTFLOW returns the face pressure at
// once; both wall types clamp (Neumann) and continue.
template <class R>
R load_pressure(const R* f, int x, int y, int z, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig& bc) {
    int c[3] = {x, y, z};
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < n[a]) continue;
        const auto& face = face_of(bc, a, c[a] < 0);
        if (periodic(bc, a) && n[a] > 0)
            c[a] = wrap(c[a], n[a]);
        else if (face.type == kOutflow)
            return static_cast<R>(face.pressure);
        else
            c[a] = c[a] < 0 ? 0 : n[a] - 1;
    }
    return f[at(c[0], c[1], c[2], nx, ny)];
}

// reference backend-cuda.cu:130-189: the first non-periodic out-of-range axis returns at
// once, from the interior value at the position clamped on every axis (earlier periodic
// axes already wrapped, later ones clamped, not wrapped).
template <class R>
R load_velocity(const R* f, int comp, int x, int y, int z, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig& bc) {
    int c[3] = {x, y, z};
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < n[a]) continue;
        const auto& face = face_of(bc, a, c[a] < 0);
        if (periodic(bc, a) && n[a] > 0) {
            c[a] = wrap(c[a], n[a]);
            continue;
        }
        int q[3];
        for (int k = 0; k < 3; ++k) q[k] = std::clamp(c[k], 0, n[k] - 1);
        const R interior = f[at(q[0], q[1], q[2], nx, ny)];
        const float prescribed = comp == 0 ? face.velocity_x : comp == 1 ? face.velocity_y : face.velocity_z;
        if (face.type == kOutflow) return interior;
        if (face.type == kFreeSlip && comp != a) return interior;
        return static_cast<R>(2) * static_cast<R>(prescribed) - interior;
    }
    return f[at(c[0], c[1], c[2], nx, ny)];
}

// periodic position wrap used by both samplers (reference backend-cuda.cu:195-206, :245-256)
template <class R>
R wrap_position(R p, int n, R h) {
    const R extent = static_cast<R>(n) * h;
    if (extent <= static_cast<R>(0)) return static_cast<R>(0);
    p = std::fmod(p, extent);
    if (p < static_cast<R>(0)) p += extent;
    return p;
}

// trilinear interpolation from 8 taps, x then y then z (reference backend-cuda.cu:208-238)
template <class R, class Tap>
R trilinear(R x, R y, R z, R h, Tap tap) {
    const R gx = x / h - static_cast<R>(0.5);
    const R gy = y / h - static_cast<R>(0.5);
    const R gz = z / h - static_cast<R>(0.5);
    const int x0 = static_cast<int>(std::floor(gx));
    const int y0 = static_cast<int>(std::floor(gy));
    const int z0 = static_cast<int>(std::floor(gz));
    const R tx = gx - static_cast<R>(x0);
    const R ty = gy - static_cast<R>(y0);
    const R tz = gz - static_cast<R>(z0);
    const R c000 = tap(x0, y0, z0), c100 = tap(x0 + 1, y0, z0);
    const R c010 = tap(x0, y0 + 1, z0), c110 = tap(x0 + 1, y0 + 1, z0);
    const R c001 = tap(x0, y0, z0 + 1), c101 = tap(x0 + 1, y0, z0 + 1);
    const R c011 = tap(x0, y0 + 1, z0 + 1), c111 = tap(x0 + 1, y0 + 1, z0 + 1);
    const R c00 = c000 + (c100 - c000) * tx;
    const R c10 = c010 + (c110 - c010) * tx;
    const R c01 = c001 + (c101 - c001) * tx;
    const R c11 = c011 + (c111 - c011) * tx;
    const R c0 = c00 + (c10 - c00) * ty;
    const R c1 = c01 + (c11 - c01) * ty;
    return c0 + (c1 - c0) * tz;
}

// reference backend-cuda.cu:191-239
template <class R>
R sample_scalar(const R* f, R x, R y, R z, int nx, int ny, int nz, R h, const StableFluidsScalarBoundaryConfig& bc) {
    if (periodic(bc, 0)) x = wrap_position(x, nx, h);
    if (periodic(bc, 1)) y = wrap_position(y, ny, h);
    if (periodic(bc, 2)) z = wrap_position(z, nz, h);
    return trilinear<R>(x, y, z, h, [&](int i, int j, int k) { return load_scalar(f, i, j, k, nx, ny, nz, bc); });
}

// reference backend-cuda.cu:241-289
template <class R>
R sample_velocity(const R* f, int comp, R x, R y, R z, int nx, int ny, int nz, R h, const StableFluidsFlowBoundaryConfig& bc) {
    if (periodic(bc, 0)) x = wrap_position(x, nx, h);
    if (periodic(bc, 1)) y = wrap_position(y, ny, h);
    if (periodic(bc, 2)) z = wrap_position(z, nz, h);
    return trilinear<R>(x, y, z, h, [&](int i, int j, int k) { return load_velocity(f, comp, i, j, k, nx, ny, nz, bc); });
}

// reference backend-cuda.cu:291-300: midpoint (RK2) backtrace
template <class R>
void trace_rk2(const R* vx, const R* vy, const R* vz, R x, R y, R z, R dt, int nx, int ny, int nz, R h, const StableFluidsFlowBoundaryConfig& bc, R out[3]) {
    const R u0 = sample_velocity(vx, 0, x, y, z, nx, ny, nz, h, bc);
    const R v0 = sample_velocity(vy, 1, x, y, z, nx, ny, nz, h, bc);
    const R w0 = sample_velocity(vz, 2, x, y, z, nx, ny, nz, h, bc);
    const R half_dt = static_cast<R>(0.5) * dt;
    const R mx = x - half_dt * u0, my = y - half_dt * v0, mz = z - half_dt * w0;
    const R u1 = sample_velocity(vx, 0, mx, my, mz, nx, ny, nz, h, bc);
    const R v1 = sample_velocity(vy, 1, mx, my, mz, nx, ny, nz, h, bc);
    const R w1 = sample_velocity(vz, 2, mx, my, mz, nx, ny, nz, h, bc);
    out[0] = x - dt * u1;
    out[1] = y - dt * v1;
    out[2] = z - dt * w1;
}

template <class R>
inline R centre(int i, R h) {
    return (static_cast<R>(i) + static_cast<R>(0.5)) * h;
}

// reference backend-cuda.cu:330-340
template <class R>
void advect_velocity(R* out, const R* src, int comp, const R* vx, const R* vy, const R* vz, R dt, int nx, int ny, int nz, R h, const StableFluidsFlowBoundaryConfig& bc) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                R p[3];
                trace_rk2(vx, vy, vz, centre(x, h), centre(y, h), centre(z, h), dt, nx, ny, nz, h, bc, p);
                out[at(x, y, z, nx, ny)] = sample_velocity(src, comp, p[0], p[1], p[2], nx, ny, nz, h, bc);
            }
}

// reference backend-cuda.cu:342-352
template <class R>
void advect_scalar(R* out, const R* src, const R* vx, const R* vy, const R* vz, R dt, int nx, int ny, int nz, R h, const StableFluidsScalarBoundaryConfig& sbc, const StableFluidsFlowBoundaryConfig& fbc) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                R p[3];
                trace_rk2(vx, vy, vz, centre(x, h), centre(y, h), centre(z, h), dt, nx, ny, nz, h, fbc, p);
                out[at(x, y, z, nx, ny)] = sample_scalar(src, p[0], p[1], p[2], nx, ny, nz, h, sbc);
            }
}

// True when an in-place half-sweep could read a same-colour value through a periodic
// wrap (odd periodic axis): then the half-sweep reads a snapshot instead.
inline bool needs_snapshot(bool px, bool py, bool pz, int nx, int ny, int nz) {
    return (px && (nx & 1)) || (py && (ny & 1)) || (pz && (nz & 1));
}

// One smoother iteration family. `Nb(src, x, y, z)` returns the 6-neighbour sum in the
// reference order (x-1, x+1, y-1, y+1, z-1, z+1); `Upd(b, nb)` the new value.
// RBGS: red ((x+y+z)&1 == 0) then black (reference backend-cuda.cu:628-640, :646-649,
// :659-662). Jacobi: every cell from the previous iterate.
template <class R, class Nb, class Upd>
void smooth(R* x, const R* b, int iterations, int smoother, int nx, int ny, int nz, bool snapshot, Nb nb, Upd upd) {
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::vector<R> prev;
    if (smoother == SFO_SMOOTHER_JACOBI || snapshot) prev.resize(n);
    for (int it = 0; it < iterations; ++it) {
        if (smoother == SFO_SMOOTHER_JACOBI) {
            std::memcpy(prev.data(), x, n * sizeof(R));
#pragma omp parallel for schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int i = 0; i < nx; ++i) {
                        const std::size_t k = at(i, y, z, nx, ny);
                        x[k] = upd(b[k], nb(prev.data(), i, y, z));
                    }
            continue;
        }
        for (int parity = 0; parity < 2; ++parity) {
            const R* src = x;
            if (snapshot) {
                std::memcpy(prev.data(), x, n * sizeof(R));
                src = prev.data();
            }
#pragma omp parallel for schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int i = (parity + y + z) & 1; i < nx; i += 2) {
                        const std::size_t k = at(i, y, z, nx, ny);
                        x[k] = upd(b[k], nb(src, i, y, z));
                    }
        }
    }
}

// reference backend-cuda.cu:366-375 (scalar diffusion, x = (b + a*sum)/(1+6a))
template <class R>
void smooth_scalar(R* x, const R* b, R alpha, int iterations, int smoother, int nx, int ny, int nz, const StableFluidsScalarBoundaryConfig& bc) {
    const R denom = static_cast<R>(1) + static_cast<R>(6) * alpha;
    smooth(x, b, iterations, smoother, nx, ny, nz, needs_snapshot(periodic(bc, 0), periodic(bc, 1), periodic(bc, 2), nx, ny, nz),
        [&](const R* f, int i, int j, int k) {
            return load_scalar(f, i - 1, j, k, nx, ny, nz, bc) + load_scalar(f, i + 1, j, k, nx, ny, nz, bc) + load_scalar(f, i, j - 1, k, nx, ny, nz, bc) + load_scalar(f, i, j + 1, k, nx, ny, nz, bc) + load_scalar(f, i, j, k - 1, nx, ny, nz, bc) + load_scalar(f, i, j, k + 1, nx, ny, nz, bc);
        },
        [&](R bv, R s) { return (bv + alpha * s) / denom; });
}

// reference backend-cuda.cu:354-364 (velocity component diffusion)
template <class R>
void smooth_velocity(R* x, const R* b, R alpha, int comp, int iterations, int smoother, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig& bc) {
    const R denom = static_cast<R>(1) + static_cast<R>(6) * alpha;
    smooth(x, b, iterations, smoother, nx, ny, nz, needs_snapshot(periodic(bc, 0), periodic(bc, 1), periodic(bc, 2), nx, ny, nz),
        [&](const R* f, int i, int j, int k) {
            return load_velocity(f, comp, i - 1, j, k, nx, ny, nz, bc) + load_velocity(f, comp, i + 1, j, k, nx, ny, nz, bc) + load_velocity(f, comp, i, j - 1, k, nx, ny, nz, bc) + load_velocity(f, comp, i, j + 1, k, nx, ny, nz, bc) + load_velocity(f, comp, i, j, k - 1, nx, ny, nz, bc)
                 + load_velocity(f, comp, i, j, k + 1, nx, ny, nz, bc);
        },
        [&](R bv, R s) { return (bv + alpha * s) / denom; });
}

// reference backend-cuda.cu:398-407 (pressure Poisson, p = (sum - h^2 div)/6)
template <class R>
void smooth_pressure(R* p, const R* div, R h2, int iterations, int smoother, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig& bc) {
    smooth(p, div, iterations, smoother, nx, ny, nz, needs_snapshot(periodic(bc, 0), periodic(bc, 1), periodic(bc, 2), nx, ny, nz),
        [&](const R* f, int i, int j, int k) {
            return load_pressure(f, i - 1, j, k, nx, ny, nz, bc) + load_pressure(f, i + 1, j, k, nx, ny, nz, bc) + load_pressure(f, i, j - 1, k, nx, ny, nz, bc) + load_pressure(f, i, j + 1, k, nx, ny, nz, bc) + load_pressure(f, i, j, k - 1, nx, ny, nz, bc) + load_pressure(f, i, j, k + 1, nx, ny, nz, bc);
        },
        [&](R dv, R s) { return (s - h2 * dv) / static_cast<R>(6); });
}

// reference backend-cuda.cu:386-396
template <class R>
void divergence(R* div, const R* vx, const R* vy, const R* vz, int nx, int ny, int nz, R h, const StableFluidsFlowBoundaryConfig& bc) {
    const R inv_2h = static_cast<R>(0.5) / h;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const R ddx = (load_velocity(vx, 0, x + 1, y, z, nx, ny, nz, bc) - load_velocity(vx, 0, x - 1, y, z, nx, ny, nz, bc)) * inv_2h;
                const R ddy = (load_velocity(vy, 1, x, y + 1, z, nx, ny, nz, bc) - load_velocity(vy, 1, x, y - 1, z, nx, ny, nz, bc)) * inv_2h;
                const R ddz = (load_velocity(vz, 2, x, y, z + 1, nx, ny, nz, bc) - load_velocity(vz, 2, x, y, z - 1, nx, ny, nz, bc)) * inv_2h;
                div[at(x, y, z, nx, ny)] = ddx + ddy + ddz;
            }
}

// reference backend-cuda.cu:409-422 (in place: v -= central gradient of p)
template <class R>
void project(R* vx, R* vy, R* vz, const R* p, int nx, int ny, int nz, R h, const StableFluidsFlowBoundaryConfig& bc) {
    const R inv_2h = static_cast<R>(0.5) / h;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const std::size_t k = at(x, y, z, nx, ny);
                const R gx = (load_pressure(p, x + 1, y, z, nx, ny, nz, bc) - load_pressure(p, x - 1, y, z, nx, ny, nz, bc)) * inv_2h;
                const R gy = (load_pressure(p, x, y + 1, z, nx, ny, nz, bc) - load_pressure(p, x, y - 1, z, nx, ny, nz, bc)) * inv_2h;
                const R gz = (load_pressure(p, x, y, z + 1, nx, ny, nz, bc) - load_pressure(p, x, y, z - 1, nx, ny, nz, bc)) * inv_2h;
                vx[k] = vx[k] - gx;
                vy[k] = vy[k] - gy;
                vz[k] = vz[k] - gz;
            }
}

struct SimBase {
    virtual ~SimBase() = default;
    virtual int update_scalar(uint32_t h, const float* v, bool source) = 0;
    virtual int update_vector(uint32_t h, const float* x, const float* y, const float* z) = 0;
    virtual int set_velocity(const float* x, const float* y, const float* z) = 0;
    virtual void step() = 0;
    virtual int get(uint32_t kind, uint32_t h, double* o0, double* o1, double* o2) = 0;
};

// Whole step, reference backend-cuda.cu:583-606 (initial state) and :612-671 (order).
template <class R>
struct Sim final : SimBase {
    StableFluidsSimulationConfig cfg{};
    int smoother = SFO_SMOOTHER_RBGS;
    int nx = 0, ny = 0, nz = 0;
    std::size_t n = 0;
    std::vector<R> vx, vy, vz, tx, ty, tz, p, div, mag;
    struct Scalar {
        StableFluidsScalarFieldDesc desc;
        std::vector<R> data, temp, source;
    };
    struct Vector {
        StableFluidsVectorFieldDesc desc;
        std::vector<R> x, y, z;
    };
    std::vector<Scalar> scalars;
    std::vector<Vector> vectors;

    Sim(const StableFluidsContextCreateDesc& d, int smoother_) : cfg(d.config), smoother(smoother_) {
        nx = cfg.nx;
        ny = cfg.ny;
        nz = cfg.nz;
        n = static_cast<std::size_t>(nx) * ny * nz;
        for (auto* a : {&vx, &vy, &vz, &tx, &ty, &tz, &p, &div, &mag}) a->assign(n, R(0));
        for (uint32_t i = 0; i < d.scalar_field_count; ++i) {
            const auto& sd = d.scalar_fields[i];
            scalars.push_back(Scalar{sd, std::vector<R>(n, static_cast<R>(sd.initial_value)), std::vector<R>(n, static_cast<R>(sd.initial_value)), std::vector<R>(n, R(0))});
        }
        for (uint32_t i = 0; i < d.vector_field_count; ++i) {
            const auto& vd = d.vector_fields[i];
            vectors.push_back(Vector{vd, std::vector<R>(n, static_cast<R>(vd.initial_value_x)), std::vector<R>(n, static_cast<R>(vd.initial_value_y)), std::vector<R>(n, static_cast<R>(vd.initial_value_z))});
        }
    }

    static void load(std::vector<R>& dst, const float* src) {
        if (!src) {
            std::fill(dst.begin(), dst.end(), R(0));
            return;
        }
        for (std::size_t i = 0; i < dst.size(); ++i) dst[i] = static_cast<R>(src[i]);
    }
    int update_scalar(uint32_t h, const float* v, bool source) override {
        if (h == 0 || h > scalars.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        load(source ? scalars[h - 1].source : scalars[h - 1].data, v);
        return 0;
    }
    int update_vector(uint32_t h, const float* x, const float* y, const float* z) override {
        if (h == 0 || h > vectors.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        load(vectors[h - 1].x, x);
        load(vectors[h - 1].y, y);
        load(vectors[h - 1].z, z);
        return 0;
    }
    int set_velocity(const float* x, const float* y, const float* z) override {
        load(vx, x);
        load(vy, y);
        load(vz, z);
        return 0;
    }

    void step() override {
        const auto& bc = cfg.flow_boundary;
        const R dt = static_cast<R>(cfg.dt);
        const R h = static_cast<R>(cfg.cell_size);
        // add force (reference :612-615, kernel :310-319)
        for (auto& v : vectors) {
            if (v.desc.usage != STABLE_FLUIDS_VECTOR_FIELD_FORCE) continue;
#pragma omp parallel for schedule(static)
            for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(n); ++i) {
                vx[i] += dt * v.x[i];
                vy[i] += dt * v.y[i];
                vz[i] += dt * v.z[i];
            }
        }
        // self-advection of the three components (reference :617-619)
        advect_velocity(tx.data(), vx.data(), 0, vx.data(), vy.data(), vz.data(), dt, nx, ny, nz, h, bc);
        advect_velocity(ty.data(), vy.data(), 1, vx.data(), vy.data(), vz.data(), dt, nx, ny, nz, h, bc);
        advect_velocity(tz.data(), vz.data(), 2, vx.data(), vy.data(), vz.data(), dt, nx, ny, nz, h, bc);
        // viscosity (reference :621-642); alpha <= 0 leaves v as forced (advection dropped)
        const R alpha = dt * static_cast<R>(cfg.viscosity) / (h * h);
        if (alpha > R(0)) {
            vx = tx;
            vy = ty;
            vz = tz;
            smooth_velocity(vx.data(), tx.data(), alpha, 0, cfg.diffuse_iterations, smoother, nx, ny, nz, bc);
            smooth_velocity(vy.data(), ty.data(), alpha, 1, cfg.diffuse_iterations, smoother, nx, ny, nz, bc);
            smooth_velocity(vz.data(), tz.data(), alpha, 2, cfg.diffuse_iterations, smoother, nx, ny, nz, bc);
        }
        // projection (reference :643-650): pressure restarts from zero every step
        divergence(div.data(), vx.data(), vy.data(), vz.data(), nx, ny, nz, h, bc);
        std::fill(p.begin(), p.end(), R(0));
        smooth_pressure(p.data(), div.data(), h * h, cfg.pressure_iterations, smoother, nx, ny, nz, bc);
        project(vx.data(), vy.data(), vz.data(), p.data(), nx, ny, nz, h, bc);
        // scalars (reference :652-668)
        for (auto& s : scalars) {
#pragma omp parallel for schedule(static)
            for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(n); ++i) s.data[i] += dt * s.source[i];
            advect_scalar(s.temp.data(), s.data.data(), vx.data(), vy.data(), vz.data(), dt, nx, ny, nz, h, s.desc.boundary, bc);
            const R alpha_s = dt * static_cast<R>(s.desc.diffusion) / (h * h);
            if (alpha_s > R(0)) {
                s.data = s.temp;
                smooth_scalar(s.data.data(), s.temp.data(), alpha_s, cfg.diffuse_iterations, smoother, nx, ny, nz, s.desc.boundary);
            }
            if (s.desc.dissipation > 0.0f) {
                const R factor = R(1) / (R(1) + dt * static_cast<R>(s.desc.dissipation));
#pragma omp parallel for schedule(static)
                for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(n); ++i) s.data[i] = s.data[i] * factor;
            }
        }
        // diagnostics (reference :670-671)
        divergence(div.data(), vx.data(), vy.data(), vz.data(), nx, ny, nz, h, bc);
#pragma omp parallel for schedule(static)
        for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(n); ++i) mag[i] = std::sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    }

    static void out(const std::vector<R>& src, double* dst) {
        if (dst)
            for (std::size_t i = 0; i < src.size(); ++i) dst[i] = static_cast<double>(src[i]);
    }
    int get(uint32_t kind, uint32_t h, double* o0, double* o1, double* o2) override {
        switch (kind) {
        case STABLE_FLUIDS_VIEW_SCALAR_FIELD_DATA:
        case STABLE_FLUIDS_VIEW_SCALAR_FIELD_SOURCE:
            if (h == 0 || h > scalars.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            out(kind == STABLE_FLUIDS_VIEW_SCALAR_FIELD_DATA ? scalars[h - 1].data : scalars[h - 1].source, o0);
            return 0;
        case STABLE_FLUIDS_VIEW_VECTOR_FIELD:
            if (h == 0 || h > vectors.size()) return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
            out(vectors[h - 1].x, o0);
            out(vectors[h - 1].y, o1);
            out(vectors[h - 1].z, o2);
            return 0;
        case STABLE_FLUIDS_VIEW_FLOW_VELOCITY:
            out(vx, o0);
            out(vy, o1);
            out(vz, o2);
            return 0;
        case STABLE_FLUIDS_VIEW_FLOW_VELOCITY_MAGNITUDE: out(mag, o0); return 0;
        case STABLE_FLUIDS_VIEW_FLOW_PRESSURE: out(p, o0); return 0;
        case STABLE_FLUIDS_VIEW_FLOW_DIVERGENCE: out(div, o0); return 0;
        default: return STABLE_FLUIDS_RESULT_BACKEND_FAILURE;
        }
    }
};

} // namespace sfo

using namespace sfo;

extern "C" {

void* sfo_create(const StableFluidsContextCreateDesc* desc, int precision, int smoother) {
    if (!desc || desc->config.nx <= 0 || desc->config.ny <= 0 || desc->config.nz <= 0) return nullptr;
    if (precision == 32) return new (std::nothrow) Sim<float>(*desc, smoother);
    return new (std::nothrow) Sim<double>(*desc, smoother);
}
void sfo_destroy(void* sim) { delete static_cast<SimBase*>(sim); }
int sfo_update_scalar_field(void* sim, uint32_t h, const float* v) { return static_cast<SimBase*>(sim)->update_scalar(h, v, false); }
int sfo_update_scalar_field_source(void* sim, uint32_t h, const float* v) { return static_cast<SimBase*>(sim)->update_scalar(h, v, true); }
int sfo_update_vector_field(void* sim, uint32_t h, const float* x, const float* y, const float* z) { return static_cast<SimBase*>(sim)->update_vector(h, x, y, z); }
int sfo_set_flow_velocity(void* sim, const float* x, const float* y, const float* z) { return static_cast<SimBase*>(sim)->set_velocity(x, y, z); }
void sfo_step(void* sim, int steps) {
    for (int i = 0; i < steps; ++i) static_cast<SimBase*>(sim)->step();
}
int sfo_get_field(void* sim, uint32_t kind, uint32_t h, double* o0, double* o1, double* o2) { return static_cast<SimBase*>(sim)->get(kind, h, o0, o1, o2); }
// Threads the next parallel loops use (n >= 1; the CPU baseline times 1 and all).
void sfo_set_threads(int n) {
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int sfo_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

double sfo_load_scalar(const double* f, int x, int y, int z, int nx, int ny, int nz, const StableFluidsScalarBoundaryConfig* bc) { return load_scalar(f, x, y, z, nx, ny, nz, *bc); }
double sfo_load_pressure(const double* f, int x, int y, int z, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig* bc) { return load_pressure(f, x, y, z, nx, ny, nz, *bc); }
double sfo_load_velocity(const double* f, int c, int x, int y, int z, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig* bc) { return load_velocity(f, c, x, y, z, nx, ny, nz, *bc); }
double sfo_sample_scalar(const double* f, double x, double y, double z, int nx, int ny, int nz, double h, const StableFluidsScalarBoundaryConfig* bc) { return sample_scalar(f, x, y, z, nx, ny, nz, h, *bc); }
double sfo_sample_velocity(const double* f, int c, double x, double y, double z, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc) { return sample_velocity(f, c, x, y, z, nx, ny, nz, h, *bc); }
void sfo_trace_rk2(const double* vx, const double* vy, const double* vz, double x, double y, double z, double dt, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc, double* out3) { trace_rk2(vx, vy, vz, x, y, z, dt, nx, ny, nz, h, *bc, out3); }
void sfo_advect_velocity(double* out, const double* src, int c, const double* vx, const double* vy, const double* vz, double dt, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc) { advect_velocity(out, src, c, vx, vy, vz, dt, nx, ny, nz, h, *bc); }
void sfo_advect_scalar(double* out, const double* src, const double* vx, const double* vy, const double* vz, double dt, int nx, int ny, int nz, double h, const StableFluidsScalarBoundaryConfig* sbc, const StableFluidsFlowBoundaryConfig* fbc) { advect_scalar(out, src, vx, vy, vz, dt, nx, ny, nz, h, *sbc, *fbc); }
void sfo_smooth_scalar(double* x, const double* b, double alpha, int it, int sm, int nx, int ny, int nz, const StableFluidsScalarBoundaryConfig* bc) { smooth_scalar(x, b, alpha, it, sm, nx, ny, nz, *bc); }
void sfo_smooth_velocity(double* x, const double* b, double alpha, int c, int it, int sm, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig* bc) { smooth_velocity(x, b, alpha, c, it, sm, nx, ny, nz, *bc); }
void sfo_smooth_pressure(double* p, const double* div, double h2, int it, int sm, int nx, int ny, int nz, const StableFluidsFlowBoundaryConfig* bc) { smooth_pressure(p, div, h2, it, sm, nx, ny, nz, *bc); }
void sfo_divergence(double* div, const double* vx, const double* vy, const double* vz, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc) { divergence(div, vx, vy, vz, nx, ny, nz, h, *bc); }
void sfo_project(double* vx, double* vy, double* vz, const double* p, int nx, int ny, int nz, double h, const StableFluidsFlowBoundaryConfig* bc) { project(vx, vy, vz, p, nx, ny, nz, h, *bc); }

} // extern "C"
