// init_spheres.h -- host side of the sphere initial condition (k_init_spheres):
// the per-block candidate lists and the launch, shared by single grids and
// ensemble members.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <vector>

#include "k_misc.h"

// The blocks along one axis of n sites (nblk blocks of SPH_B) that hold a
// site within periodic distance w of coordinate cen, as flags in `hit`.
static inline void sph_axis_blocks(double cen, double w, int n, std::vector<char> &hit)
{
    const int nblk = (n + SPH_B - 1) / SPH_B;
    hit.assign((size_t)nblk, 0);
    if (2.0 * w + 2.0 >= n) {
        hit.assign((size_t)nblk, 1);
        return;
    }
    const long lo = (long)floor(cen - w), hi = (long)ceil(cen + w);
    for (long p = lo; p <= hi; p++) {
        long q = p % n;
        if (q < 0)
            q += n;
        hit[(size_t)(q / SPH_B)] = 1;
    }
}

// Fill the device arrays (s, l, c) of one nz x ny x nx grid with the sphere
// initial condition, on `stream`, and wait for it.  centers_zyx: n (z, y, x)
// triples in cell units; radii: n values.  Returns the first HIP error, or
// hipErrorInvalidValue when the candidate lists would not fit 32-bit
// indices.
template <typename T>
static hipError_t init_spheres_device(hipStream_t stream, T *s, T *l, T *c, int nx, int ny, int nz,
                                      int n, const double *centers_zyx, const double *radii,
                                      double c_liq, double c_span)
{
    const int nbx = (nx + SPH_B - 1) / SPH_B, nby = (ny + SPH_B - 1) / SPH_B,
              nbz = (nz + SPH_B - 1) / SPH_B;
    const size_t nb = (size_t)nbx * nby * nbz;
    std::vector<int> off(nb + 1, 0);
    std::vector<char> hz, hy, hx;
    std::vector<double> rec(4 * (size_t)(n > 0 ? n : 1), 0.0);
    // pass 1: count the candidates of every block
    std::vector<int> bz, by, bx;
    auto axis_list = [](const std::vector<char> &h, std::vector<int> &v) {
        v.clear();
        for (size_t i = 0; i < h.size(); i++)
            if (h[i])
                v.push_back((int)i);
    };
    for (int pass = 0; pass < 2; pass++) {
        std::vector<int> fill;
        std::vector<int> list;
        if (pass == 1) {
            for (size_t b = 0; b < nb; b++)
                off[b + 1] += off[b];
            if ((uint64_t)off[nb] > (uint64_t)INT32_MAX)
                return hipErrorInvalidValue;
            list.resize((size_t)off[nb] > 0 ? (size_t)off[nb] : 1);
            fill.assign(off.begin(), off.end() - 1);
        }
        for (int i = 0; i < n; i++) {
            const double w = radii[i] + SPH_CUT;
            sph_axis_blocks(centers_zyx[3 * i + 0], w, nz, hz);
            sph_axis_blocks(centers_zyx[3 * i + 1], w, ny, hy);
            sph_axis_blocks(centers_zyx[3 * i + 2], w, nx, hx);
            axis_list(hz, bz);
            axis_list(hy, by);
            axis_list(hx, bx);
            for (int zb : bz)
                for (int yb : by)
                    for (int xb : bx) {
                        const size_t b = ((size_t)zb * nby + yb) * nbx + xb;
                        if (pass == 0) {
                            if (off[b + 1] == INT32_MAX)
                                return hipErrorInvalidValue;
                            off[b + 1]++;
                        } else {
                            list[(size_t)fill[b]++] = i;
                        }
                    }
            if (pass == 0) {
                rec[4 * (size_t)i + 0] = centers_zyx[3 * i + 0];
                rec[4 * (size_t)i + 1] = centers_zyx[3 * i + 1];
                rec[4 * (size_t)i + 2] = centers_zyx[3 * i + 2];
                rec[4 * (size_t)i + 3] = radii[i];
            }
        }
        if (pass == 1) {
            double *dsph = nullptr;
            int *doff = nullptr, *dlist = nullptr;
            hipError_t e = hipMalloc(&dsph, rec.size() * sizeof(double));
            if (e == hipSuccess)
                e = hipMalloc(&doff, off.size() * sizeof(int));
            if (e == hipSuccess)
                e = hipMalloc(&dlist, list.size() * sizeof(int));
            if (e == hipSuccess)
                e = hipMemcpyAsync(dsph, rec.data(), rec.size() * sizeof(double),
                                   hipMemcpyHostToDevice, stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(doff, off.data(), off.size() * sizeof(int),
                                   hipMemcpyHostToDevice, stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dlist, list.data(), list.size() * sizeof(int),
                                   hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_init_spheres<T>, dim3((unsigned)nb), dim3(SPH_B * SPH_B), 0,
                                   stream, s, l, c, nx, ny, nz, (const double *)dsph,
                                   (const int *)doff, (const int *)dlist, nbx, nby, n > 0 ? 1 : 0,
                                   c_liq, c_span);
                e = hipGetLastError();
            }
            const hipError_t e2 = hipStreamSynchronize(stream);
            if (e == hipSuccess)
                e = e2;
            if (dsph)
                hipFree(dsph);
            if (doff)
                hipFree(doff);
            if (dlist)
                hipFree(dlist);
            return e;
        }
    }
    return hipSuccess;
}
