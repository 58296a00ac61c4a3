// sf_warm.hpp -- learned pressure warm start (guess, residual test, fallback flag) and the
// FP64 conjugate-gradient pressure solve that makes its training pairs (sf_warm.hip).
//
// The pressure system of a step is A p = b with (A p)_c = 6 p_c - sum of the 6
// neighbours, ghosts by the pressure rules (walls copy the cell, outflow faces give
// their pressure, periodic axes wrap), b = -h^2 div (reference backend-cuda.cu:374-384).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sf_device.hpp"

namespace sf::warm {

constexpr int kPartialBlocks = 8192; // enough waves in flight for the grid-stride stages

// Coarse network input in2[v][2] = (mean of the f^3 fine div values, wall mask) on the
// coarse grid cd (fine = f * cd), and the per-block sums of squares of the means in
// partials (restrict_partials(cd) of them; unet::launch_pack_input_sigma turns them into
// sigma = RMS of the coarse divergence). partials: kPartialBlocks doubles.
hipError_t launch_restrict_input(hipStream_t s, const float* div, float* in2, double* partials, Dims cd, int f, const StencilBC& p_bc);
unsigned restrict_partials(Dims cd);
// p0[fine] = scale * sigma * trilinear(q)[fine] (cell-centred, clamped at the edges),
// scale = (f h)^2, sigma = scal[0]. The coarse grid is at most 2048 cells wide.
hipError_t launch_prolong(hipStream_t s, const float* q, const float* scal, float scale, float* p0, Dims cd, int f);
// Residual test of the guess p0: g = ||b - A p0||, z = ||b|| (= residual of the zero
// guess), FP64. stats[0] = g, stats[1] = z, stats[2] = 1 when g > ratio * z (fallback).
// skip[0] = 0 on fallback (the fallback passes run), 1 otherwise; skip[1] = 0 (the grid-
// barrier count of a one-launch fallback solve); on fallback p0 is zeroed when
// zero_on_fallback. partials: 2 * kPartialBlocks doubles.
hipError_t launch_guess_test(hipStream_t s, float* p0, const float* div, float h2, double ratio, double* partials, double* stats, int* skip, Dims d, const StencilBC& p_bc, bool zero_on_fallback);

// FP64 conjugate gradients on A p = b (b from div as above); x holds the start value
// and receives the solution. Pure Neumann/periodic systems (no outflow face) are
// singular: b is projected to zero mean and the solution returned with zero mean.
// Scratch: r, p, ap (n doubles each), partials (2 * kPartialBlocks doubles), scal
// (8 doubles). Synchronises on `s` every `check` iterations to test ||r|| <= tol ||b||.
struct CgResult {
    int iterations;
    double rel_residual;
};
hipError_t cg_solve(hipStream_t s, const float* div, float h2, double* x, double* r, double* p, double* ap, double* partials, double* scal, Dims d, const StencilBC& p_bc, double tol, int max_iter, CgResult* out);
// Training pair on the grid coarsened f times from the divergence and the solution x:
// in2[v] = (div_c / sigma, wall mask), target[v] = x_c / ((f h)^2 sigma), where _c is
// the mean of the f^3 fine cells and sigma = RMS(div_c).
hipError_t launch_pair(hipStream_t s, const float* div, const double* x, float h2, float* in2, float* target, double* partials, double* scal, Dims d, int f, const StencilBC& p_bc);

} // namespace sf::warm
