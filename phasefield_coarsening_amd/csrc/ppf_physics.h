// ppf_physics.h -- the site arithmetic of one explicit Moelans step: the
// FP64 form that reproduces the reference bit for bit, and an FP32 form.
//
// Operand order follows phasefield.c:324-384 term by term (C evaluates left
// to right), and the library is compiled with -ffp-contract=off, so each
// product and sum is rounded on its own exactly as the reference's x86-64
// build rounds it.  FP64 division on gfx950 is the correctly rounded IEEE
// sequence, so the results are bitwise equal to the reference's.
//
// Neighbour order everywhere is the reference's: i+1, i-1, j+1, j-1, k+1,
// k-1 = z+1, z-1, y+1, y-1, x+1, x-1.
#pragma once

#include <hip/hip_runtime.h>

// Constants hoisted out of the loops, each computed on the host with the
// expression the reference uses (phasefield.c:322-323, 337-339, 351-358).
struct KConst {
    double AS, AL;
    double dA;     // AL - AS
    double tmp5;   // AS*AL
    double tmp6;   // cL0 - cS0
    double cL0, cS0;
    double ML;
    double dM;     // MS - ML
    double tmp1;   // -1.0*L*dt
    double tmp2;   // AS*cS0 - AL*cL0
    double tmp3;   // dt*0.5*h2
    double hAS;    // 0.5*AS
    double hAL;    // 0.5*AL
    double CS, CL;
    double m;
    double g2;     // 2.0*gamma
    double kh2;    // kappa*h2
    // With AL == AS the denominator AS + (AL-AS)*h is AS + (+-0) = AS for
    // every finite h, so tmp5/(AS+(AL-AS)*h) and 1.0/(AS+(AL-AS)*h) are the
    // constants below, bit for bit (a NaN h still propagates through the
    // other factor).  The kernels use them when AL == AS (EQA = true).
    double K5;     // tmp5/AS
    double invA;   // 1.0/AS
};

// Pass A at one site (phasefield.c:327-333): hS, mu, M.
template <bool EQA>
__device__ __forceinline__ void pointwise(double s, double l, double c, const KConst &k,
                                          double &h, double &mu, double &M)
{
    h = s * s / (s * s + l * l);
    if (EQA)
        mu = k.K5 * (c - k.cL0 + k.tmp6 * h);
    else
        mu = k.tmp5 / (k.AS + k.dA * h) * (c - k.cL0 + k.tmp6 * h);
    M = k.ML + k.dM * h;
}

// Pass B + C at one site (phasefield.c:351-384, 396-398).
//   sn, ln   : etaS, etaL at the six neighbours, reference order
//   mun, Mn  : mu, M at the six neighbours, reference order
// Returns the new (s, l, c) and the increments.
template <bool EQA>
__device__ __forceinline__ void site_update(double s, double l, double c, double h, double mu,
                                            double M, const double sn[6], const double ln[6],
                                            const double mun[6], const double Mn[6],
                                            const KConst &k, double &s_new, double &l_new,
                                            double &c_new, double &ds, double &dl, double &dc)
{
    const double tmp4 = EQA ? k.invA : 1.0 / (k.AS + k.dA * h);
    const double cS = tmp4 * (k.AL * c + (1.0 - h) * k.tmp2);
    const double cL = tmp4 * (k.AS * c - h * k.tmp2);
    const double fS = k.hAS * (cS - k.cS0) * (cS - k.cS0) + k.CS;
    const double fL = k.hAL * (cL - k.cL0) * (cL - k.cL0) + k.CL;
    const double dmu = 2.0 * (s * l) / ((s * s + l * l) * (s * s + l * l)) *
                       (fS - fL - mu * (cS - cL));
    ds = k.tmp1 * (k.m * s * (s * s - 1.0 + k.g2 * l * l)
                   - k.kh2 * (sn[0] + sn[1] + sn[2] + sn[3] + sn[4] + sn[5] - 6.0 * s)
                   + l * dmu);
    dl = k.tmp1 * (k.m * l * (l * l - 1.0 + k.g2 * s * s)
                   - k.kh2 * (ln[0] + ln[1] + ln[2] + ln[3] + ln[4] + ln[5] - 6.0 * l)
                   - s * dmu);
    dc = k.tmp3 * (Mn[0] * (mun[0] - mu) -
                   Mn[1] * (mu - mun[1]) +
                   Mn[2] * (mun[2] - mu) -
                   Mn[3] * (mu - mun[3]) +
                   Mn[4] * (mun[4] - mu) -
                   Mn[5] * (mu - mun[5]) +
                   M * (mun[0] + mun[1] + mun[2] + mun[3] + mun[4] + mun[5] - 6.0 * mu));
    s_new = s + ds;
    l_new = l + dl;
    c_new = c + dc;
}

// c-only update (phasefield.c:451-461, 470): eta frozen.
__device__ __forceinline__ double dc_only(double mu, double M, const double mun[6],
                                          const double Mn[6], const KConst &k)
{
    return k.tmp3 * (Mn[0] * (mun[0] - mu) -
                     Mn[1] * (mu - mun[1]) +
                     Mn[2] * (mun[2] - mu) -
                     Mn[3] * (mu - mun[3]) +
                     Mn[4] * (mun[4] - mu) -
                     Mn[5] * (mu - mun[5]) +
                     M * (mun[0] + mun[1] + mun[2] + mun[3] + mun[4] + mun[5] - 6.0 * mu));
}

// Ghost-cell index map of one axis.  Periodic (phasefield.c:1117-1146):
// ghost 0 <- N, N+1 <- 1, i.e. -1 -> n-1 and n -> 0.  No-flux
// (phasefield.c:1158-1187): ghost 0 <- 1, N+1 <- N, i.e. -1 -> 0 and
// n -> n-1.  Only a one-cell excursion is ever needed.
template <int BC>
__device__ __forceinline__ int ghost(int i, int n)
{
    if (BC == 0)
        return i < 0 ? i + n : (i >= n ? i - n : i);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// ---------------------------------------------------------------------------
// Physics policies for the tiled kernel.  Each provides the storage type T,
// the per-site pass A (point), the stencil update (update) and the c-only
// update (conly).
// ---------------------------------------------------------------------------

// FP64, bitwise equal to the reference (the functions above).
template <bool EQA>
struct PhysF64 {
    using T = double;
    using K = KConst;
    // the two-step march keeps two planes of state per column (a third set
    // of registers would make the three-wave kernel spill)
    static constexpr bool kRotate = false;
    static __device__ __forceinline__ void point(T s, T l, T c, const K &k, T &h, T &mu, T &M)
    {
        pointwise<EQA>(s, l, c, k, h, mu, M);
    }
    static __device__ __forceinline__ void update(T s, T l, T c, T h, T mu, T M, const T sn[6],
                                                  const T ln[6], const T mun[6], const T Mn[6],
                                                  const K &k, T &so, T &lo, T &co, double &ds,
                                                  double &dl, double &dc)
    {
        site_update<EQA>(s, l, c, h, mu, M, sn, ln, mun, Mn, k, so, lo, co, ds, dl, dc);
    }
    static __device__ __forceinline__ T conly(T mu, T M, const T mun[6], const T Mn[6],
                                              const K &k)
    {
        return dc_only(mu, M, mun, Mn, k);
    }
};

// Collapsed form for AS == AL (the FP32 policy and the fast FP64 policy).
// With equal free-energy curvatures the chemistry collapses
// (phasefield.c:351-359): cS - cL = cS0 - cL0, fS - fL = CS - CL and
// mu = A*u with the dimensionless u = c - cL0 + (cL0 - cS0)*h.  Writing
// M = ML*mm with mm = 1 + ((MS - ML)/ML)*h and r = 1/(s^2 + l^2), one step is
//   h  = s^2 r
//   D  = s l r^2 (b0 + b1*u)
//   ds = a1*s*(s^2 - 1 + 2 gamma l^2) + a2*Lap6(s) + l*D
//   dl = a1*l*(l^2 - 1 + 2 gamma s^2) + a2*Lap6(l) - s*D
//   dc = d0 * sum_nb (mm_nb + mm)*(u_nb - u)            (flux form)
// with a1 = tmp1*m, a2 = -tmp1*kappa*h2, b0 = 2*tmp1*(CS - CL),
// b1 = 2*tmp1*A*(cL0 - cS0), d0 = tmp3*ML*A, tmp1 = -L*dt, tmp3 = dt*h2/2
// (the factor 2 of 2 s l/(s^2 + l^2)^2, phasefield.c:359, is in b0 and b1).
// Every quantity is O(1) or a small increment: no 1e8-sized free energies
// are subtracted (the reference subtracts fS and fL, each ~1e8, whose
// difference is the constant CS - CL).  The flux (mm_i + mm_j)*(u_j - u_i)
// is computed identically, up to sign, at both ends of a face, so the grid
// sum of dc differs from zero only by the rounding of each site's six-term
// sum and of the final products.
//
// Both policies evaluate it with fused multiply-adds and one reciprocal per
// site (shared by h and D), instead of the reference's two divides.  The
// point() slot that carries h in PhysF64 carries r here: the update needs
// r for D, and nothing needs h of a site once its u and mm are known.
template <typename F>
struct KConstC {
    F cL0, tmp6, rM, a1, a2, g2, b0, b1, d0;
};
using KConst32 = KConstC<float>;
using KConstF = KConstC<double>;

// 1/x: the hardware reciprocal refined by Newton-Raphson steps
// r <- r + r (1 - x r), each squaring the relative error.  FP64: v_rcp_f64
// is good to about 2^-26 (two steps reach the last bit); FP32: v_rcp_f32 is
// within 1 ulp and is used as it is.  x = s^2 + l^2 is O(1) on every state
// the scheme produces (s + l ~ 1); x = 0 (s = l = 0) gives NaN, as 0/0 does
// in the reference's h.
__device__ __forceinline__ double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename F>
struct PhysC {
    using T = F;
    using K = KConstC<F>;
    // the two-step march rotates three register sets per column (no copies)
    static constexpr bool kRotate = true;
    static __device__ __forceinline__ void point(T s, T l, T c, const K &k, T &r, T &u, T &mm)
    {
        const T ss = s * s;
        r = recip(fma(l, l, ss));
        const T h = ss * r;
        u = fma(k.tmp6, h, c - k.cL0);
        mm = fma(k.rM, h, (T)1);
    }
    static __device__ __forceinline__ void update(T s, T l, T c, T r, T u, T mm, const T sn[6],
                                                  const T ln[6], const T un[6], const T mn[6],
                                                  const K &k, T &so, T &lo, T &co, double &ds_out,
                                                  double &dl_out, double &dc_out)
    {
        const T D = (s * l) * (r * r) * fma(k.b1, u, k.b0);
        const T lap_s = fma((T)-6, s, ((sn[0] + sn[1]) + (sn[2] + sn[3])) + (sn[4] + sn[5]));
        const T lap_l = fma((T)-6, l, ((ln[0] + ln[1]) + (ln[2] + ln[3])) + (ln[4] + ln[5]));
        const T ds = fma(k.a1 * s, fma(k.g2 * l, l, fma(s, s, (T)-1)), fma(k.a2, lap_s, l * D));
        const T dl = fma(k.a1 * l, fma(k.g2 * s, s, fma(l, l, (T)-1)), fma(k.a2, lap_l, -(s * D)));
        const T dc = conly(u, mm, un, mn, k);
        so = s + ds;
        lo = l + dl;
        co = c + dc;
        ds_out = ds;
        dl_out = dl;
        dc_out = dc;
    }
    static __device__ __forceinline__ T conly(T u, T mm, const T un[6], const T mn[6], const K &k)
    {
        T f = (mn[0] + mm) * (un[0] - u);
        f = fma(mn[1] + mm, un[1] - u, f);
        T g = (mn[2] + mm) * (un[2] - u);
        g = fma(mn[3] + mm, un[3] - u, g);
        T q = (mn[4] + mm) * (un[4] - u);
        q = fma(mn[5] + mm, un[5] - u, q);
        return k.d0 * ((f + g) + q);
    }
};

// FP32 form: fields stored and stepped in single precision.
using PhysF32 = PhysC<float>;
// Fast FP64 form: the collapsed arithmetic in double precision.  Not bitwise
// equal to the reference (different rounding sequence), within 1e-10 per
// field of it after 1000 steps.
using PhysF64F = PhysC<double>;
