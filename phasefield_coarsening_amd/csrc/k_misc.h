// k_misc.h -- state statistics, the sphere initial condition and small
// helpers.  None of these is on the per-step path.
#pragma once

#include "ppf_physics.h"
#include "ppf_reduce.h"

// Per-workgroup partial statistics of the state.  Record per block:
//   [0,1] sum c as double-double, [2] sum h, [3] count etaS > 0.5,
//   [4] sum etaS, [5] min c, [6] max c, [7] non-finite sites,
//   [8] sum h (1 - h), an interface measure.
constexpr int STAT_W = 9;

template <typename T>
__global__ __launch_bounds__(256) void k_stats(const T *__restrict__ s, const T *__restrict__ l,
                                               const T *__restrict__ c, size_t n,
                                               double *__restrict__ partials)
{
    DD sc = {0.0, 0.0};
    double sh = 0.0, ns = 0.0, ss = 0.0, mn = INFINITY, mx = -INFINITY, nf = 0.0, si = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const double sv = (double)s[q], lv = (double)l[q], cv = (double)c[q];
        if (!isfinite(sv) || !isfinite(lv) || !isfinite(cv)) {
            nf += 1.0;
            continue;
        }
        sc = dd_add1(sc, cv);
        const double hv = sv * sv / (sv * sv + lv * lv);
        sh += hv;
        si += hv * (1.0 - hv);
        ns += sv > 0.5 ? 1.0 : 0.0;
        ss += sv;
        mn = fmin(mn, cv);
        mx = fmax(mx, cv);
    }
    sc = wave_dd_sum(sc);
    sh = wave_sum(sh);
    ns = wave_sum(ns);
    ss = wave_sum(ss);
    mn = wave_min(mn);
    mx = wave_max(mx);
    nf = wave_sum(nf);
    si = wave_sum(si);
    __shared__ double part[STAT_W][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = sc.hi;
        part[1][w] = sc.lo;
        part[2][w] = sh;
        part[3][w] = ns;
        part[4][w] = ss;
        part[5][w] = mn;
        part[6][w] = mx;
        part[7][w] = nf;
        part[8][w] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t = {part[0][0], part[1][0]};
        double a = part[2][0], b = part[3][0], d = part[4][0], e = part[5][0], f = part[6][0],
               g = part[7][0], h = part[8][0];
        for (int j = 1; j < 4; j++) {
            t = dd_add(t, DD{part[0][j], part[1][j]});
            a += part[2][j];
            b += part[3][j];
            d += part[4][j];
            e = fmin(e, part[5][j]);
            f = fmax(f, part[6][j]);
            g += part[7][j];
            h += part[8][j];
        }
        double *o = partials + (size_t)blockIdx.x * STAT_W;
        o[0] = t.hi;
        o[1] = t.lo;
        o[2] = a;
        o[3] = b;
        o[4] = d;
        o[5] = e;
        o[6] = f;
        o[7] = g;
        o[8] = h;
    }
}

// Combine nb stat records (one workgroup of 256, fixed order) into out[0..7].
static __global__ __launch_bounds__(256) void k_stats_final(const double *__restrict__ p, int nb,
                                                     double *__restrict__ out)
{
    DD sc = {0.0, 0.0};
    double sh = 0.0, ns = 0.0, ss = 0.0, mn = INFINITY, mx = -INFINITY, nf = 0.0, si = 0.0;
    for (int j = threadIdx.x; j < nb; j += 256) {
        const double *r = p + (size_t)j * STAT_W;
        sc = dd_add(sc, DD{r[0], r[1]});
        sh += r[2];
        ns += r[3];
        ss += r[4];
        mn = fmin(mn, r[5]);
        mx = fmax(mx, r[6]);
        nf += r[7];
        si += r[8];
    }
    sc = wave_dd_sum(sc);
    sh = wave_sum(sh);
    ns = wave_sum(ns);
    ss = wave_sum(ss);
    mn = wave_min(mn);
    mx = wave_max(mx);
    nf = wave_sum(nf);
    si = wave_sum(si);
    __shared__ double part[STAT_W][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = sc.hi;
        part[1][w] = sc.lo;
        part[2][w] = sh;
        part[3][w] = ns;
        part[4][w] = ss;
        part[5][w] = mn;
        part[6][w] = mx;
        part[7][w] = nf;
        part[8][w] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t = {part[0][0], part[1][0]};
        double a = part[2][0], b = part[3][0], d = part[4][0], e = part[5][0], f = part[6][0],
               g = part[7][0], h = part[8][0];
        for (int j = 1; j < 4; j++) {
            t = dd_add(t, DD{part[0][j], part[1][j]});
            a += part[2][j];
            b += part[3][j];
            d += part[4][j];
            e = fmin(e, part[5][j]);
            f = fmax(f, part[6][j]);
            g += part[7][j];
            h += part[8][j];
        }
        out[0] = t.hi + t.lo;
        out[1] = a;
        out[2] = b;
        out[3] = d;
        out[4] = e;
        out[5] = f;
        out[6] = g;
        out[7] = h;
    }
}

// Sum nb triples (one workgroup of 256, fixed order) into out[0..2].
static __global__ __launch_bounds__(256) void k_sum3_final(const double *__restrict__ p, int nb,
                                                    double *__restrict__ out)
{
    double a[3] = {0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < nb; j += 256)
        for (int i = 0; i < 3; i++)
            a[i] += p[3 * (size_t)j + i];
    block_sum3<256>(a, out);
}

// Sphere initial condition.  sph holds n records (cz, cy, cx, R).  The
// profile 0.5*(1 - tanh(d/2)) falls with d = r - R, so the max over spheres
// is the profile at the smallest d.  The grid is cut into blocks of
// SPH_B^3 sites, one workgroup each (a thread per (x, y) column of the
// block, SPH_B planes of z); list[off[b] .. off[b+1]) holds the spheres
// whose periodic bounding box, widened by SPH_CUT, touches block b.  A
// site farther than SPH_CUT from every sphere surface has d > SPH_CUT
// whether or not its nearest sphere is listed, and tanh(d/2) is then 1.0
// exactly in double precision (1 - tanh(30) ~ 2e-26), so the profile is
// 0 either way: culling changes no bit of the result.
constexpr int SPH_B = 16;
constexpr double SPH_CUT = 60.0;

template <typename T>
__global__ __launch_bounds__(SPH_B *SPH_B) void k_init_spheres(
    T *__restrict__ s, T *__restrict__ l, T *__restrict__ c, int nx, int ny, int nz,
    const double *__restrict__ sph, const int *__restrict__ off, const int *__restrict__ list,
    int nbx, int nby, int any, double c_liq, double c_span)
{
    const int b = blockIdx.x;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int x = bx * SPH_B + (int)(threadIdx.x % SPH_B);
    const int y = by * SPH_B + (int)(threadIdx.x / SPH_B);
    const int z0 = bz * SPH_B;
    const int nzb = min(SPH_B, nz - z0);
    if (x >= nx || y >= ny)
        return;
    double dmin[SPH_B];
#pragma unroll
    for (int k = 0; k < SPH_B; k++)
        dmin[k] = 1e30;
    for (int j = off[b]; j < off[b + 1]; j++) {
        const double *p = sph + 4 * (size_t)list[j];
        double dy = fabs(y - p[1]);
        double dx = fabs(x - p[2]);
        dy = fmin(dy, ny - dy);
        dx = fmin(dx, nx - dx);
        const double cz = p[0], R = p[3];
#pragma unroll
        for (int k = 0; k < SPH_B; k++) {
            double dz = fabs(z0 + k - cz);
            dz = fmin(dz, nz - dz);
            const double d = sqrt(dz * dz + dy * dy + dx * dx) - R;
            dmin[k] = fmin(dmin[k], d);
        }
    }
    const size_t plane = (size_t)nx * ny;
#pragma unroll
    for (int k = 0; k < SPH_B; k++) {
        if (k < nzb) {
            const double sv = any ? 0.5 * (1.0 - tanh(dmin[k] / 2.0)) : 0.0;
            const size_t q = (size_t)(z0 + k) * plane + (size_t)y * nx + x;
            s[q] = (T)sv;
            l[q] = (T)(1.0 - sv);
            c[q] = (T)(c_liq + c_span * sv);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ p, size_t n, T v)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        p[q] = v;
}

// Double-double sum of a field: one (hi, lo) pair per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_sum_dd(const T *__restrict__ x, size_t n,
                                                double *__restrict__ partials)
{
    DD acc = {0.0, 0.0};
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        acc = dd_add1(acc, (double)x[q]);
    acc = wave_dd_sum(acc);
    __shared__ double part[2][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = acc.hi;
        part[1][w] = acc.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t = {part[0][0], part[1][0]};
        for (int j = 1; j < 4; j++)
            t = dd_add(t, DD{part[0][j], part[1][j]});
        partials[2 * (size_t)blockIdx.x] = t.hi;
        partials[2 * (size_t)blockIdx.x + 1] = t.lo;
    }
}

// Combine nb (hi, lo) pairs; out[1] = the sum, out[0] = (target - sum)/n,
// the per-site shift that puts sum(c) back on the FP64 target.  Run by one
// workgroup of 256 threads.
__device__ __forceinline__ void k_mass_delta_body(const double *__restrict__ p, int nb,
                                                  double target, double n,
                                                  double *__restrict__ out)
{
    DD acc = {0.0, 0.0};
    for (int j = threadIdx.x; j < nb; j += 256)
        acc = dd_add(acc, DD{p[2 * (size_t)j], p[2 * (size_t)j + 1]});
    acc = wave_dd_sum(acc);
    __shared__ double part[2][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = acc.hi;
        part[1][w] = acc.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t = {part[0][0], part[1][0]};
        for (int j = 1; j < 4; j++)
            t = dd_add(t, DD{part[0][j], part[1][j]});
        const double sum = t.hi + t.lo;
        // target - sum in double-double, then divided by n
        double d, e;
        two_sum(target, -t.hi, d, e);
        out[0] = ((d + (e - t.lo))) / n;
        out[1] = sum;
    }
}

static __global__ __launch_bounds__(256) void k_mass_delta(const double *__restrict__ p, int nb,
                                                    double target, double n,
                                                    double *__restrict__ out)
{
    k_mass_delta_body(p, nb, target, n, out);
}

// 64-bit mix (SplitMix64 finaliser) for counter-based random numbers.
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// c += d with stochastic rounding to FP32: the exact value t = c + d lies
// between two adjacent floats lo <= t <= hi; pick hi with probability
// (t - lo)/(hi - lo) from a hash of (site, seed).  The expected new sum is
// exactly sum(c) + n*d, so a shift far below one ulp per site still moves
// the total (round-to-nearest would drop it).  Sites start+threadIdx.x,
// then every stride.
__device__ __forceinline__ void mass_shift_body(float *__restrict__ c, size_t n, double d,
                                                uint64_t seed, size_t start, size_t stride)
{
    if (d == 0.0 || !isfinite(d))
        return;
    for (size_t q = start + threadIdx.x; q < n; q += stride) {
        const double t = (double)c[q] + d;
        const float f = (float)t;
        const double fd = (double)f;
        if (fd == t) {
            c[q] = f;
            continue;
        }
        float lo, hi;
        if (fd < t) {
            lo = f;
            hi = nextafterf(f, INFINITY);
        } else {
            hi = f;
            lo = nextafterf(f, -INFINITY);
        }
        const double p = (t - (double)lo) / ((double)hi - (double)lo);
        const double r =
            (double)(mix64(q * 0x9E3779B97F4A7C15ull + seed * 0xD1B54A32D192ED03ull + 1) >> 11) *
            0x1.0p-53;
        c[q] = r < p ? hi : lo;
    }
}

static __global__ __launch_bounds__(256) void k_mass_shift(float *__restrict__ c, size_t n,
                                                    const double *__restrict__ delta,
                                                    uint64_t seed)
{
    mass_shift_body(c, n, delta[0], seed, (size_t)blockIdx.x * 256, (size_t)gridDim.x * 256);
}

// Round or widen a field: dst[q] = (Td)src[q].
template <typename Ts, typename Td>
__global__ __launch_bounds__(256) void k_convert(const Ts *__restrict__ src, Td *__restrict__ dst,
                                                 size_t n)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        dst[q] = (Td)src[q];
}

// Per-workgroup max |a - b| and sum (a - b)^2 of one field pair, in FP64:
// partials[2 b] = max, partials[2 b + 1] = sum of squares.  A NaN difference
// makes the max NaN (fmax would drop it).
template <typename Ta, typename Tb>
__global__ __launch_bounds__(256) void k_diff(const Ta *__restrict__ a, const Tb *__restrict__ b,
                                              size_t n, double *__restrict__ partials)
{
    double mx = 0.0, ss = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const double d = (double)a[q] - (double)b[q];
        const double ad = fabs(d);
        mx = (ad > mx || ad != ad) ? ad : mx;
        ss += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double m2 = __shfl_xor(mx, o);
        mx = (m2 > mx || m2 != m2) ? m2 : mx;
        ss += __shfl_xor(ss, o);
    }
    __shared__ double part[2][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = mx;
        part[1][w] = ss;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0][0], t = part[1][0];
        for (int j = 1; j < 4; j++) {
            m = (part[0][j] > m || part[0][j] != part[0][j]) ? part[0][j] : m;
            t += part[1][j];
        }
        partials[2 * (size_t)blockIdx.x] = m;
        partials[2 * (size_t)blockIdx.x + 1] = t;
    }
}
