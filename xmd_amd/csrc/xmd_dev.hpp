// xmd_dev.hpp -- per-item physics of the step path as __host__ __device__
// functions, templated on the force precision T (double, or float in the
// mixed-precision form).  Positions and displacements are always double;
// mixed kernels narrow FP64-differenced displacements to float.
//
// Every function restates one reference functor body (file:line at each),
// with the engine's decisions D1-D30 (listed in xmd_oracle.cpp's header).
// Expression order follows the reference; the file is compiled with
// -ffp-contract=off so the FP64 decision code (lists, QA) evaluates exactly
// as written.
#pragma once
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "bm_math.hpp"

namespace xmd_dev {

template <typename T> struct v3 {
    T x, y, z;
};
template <typename T> XMD_HD inline v3<T> mk3(T x, T y, T z) { return v3<T>{x, y, z}; }
template <typename T> XMD_HD inline v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> XMD_HD inline v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> XMD_HD inline v3<T> operator-(v3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> XMD_HD inline v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> XMD_HD inline v3<T> operator*(T s, v3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> XMD_HD inline v3<T> operator/(v3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> XMD_HD inline T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> XMD_HD inline T norm_sq(v3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// FP32 reciprocal, square root and reciprocal square root: the hardware
// approximations on the device (v_rcp_f32, v_sqrt_f32, v_rsq_f32, about
// 1 ulp; every argument here is a positive normal number), correctly
// rounded on the host.  The FP64 form keeps correctly rounded operations.
XMD_HD inline float rcp_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
XMD_HD inline float sqrt_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
XMD_HD inline float rsq_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
// FP32 e^x as 2^(x log2 e) through v_exp_f32 (about 1 ulp, plus the
// rounding of the product: relative error below 3e-7 for |x| < 4).
XMD_HD inline float exp_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
#else
    return exp2f(x * 1.4426950408889634f);
#endif
}
template <typename T> XMD_HD inline T expT(T x) {
    if constexpr (std::is_same<T, float>::value) return exp_f(x);
    else return exp(x);
}
template <typename T> XMD_HD inline T rcpT(T x) {
    if constexpr (std::is_same<T, float>::value) return rcp_f(x);
    else return T(1.0) / x;
}
template <typename T> XMD_HD inline T sqrtT(T x) {
    if constexpr (std::is_same<T, float>::value) return sqrt_f(x);
    else return sqrt(x);
}
template <typename T> XMD_HD inline T norm(v3<T> a) { return sqrtT(norm_sq(a)); }
template <typename T> XMD_HD inline T norm_inv(v3<T> a) {
    if constexpr (std::is_same<T, float>::value) return rsq_f(norm_sq(a));
    else return T(1.0) / norm(a);
}
template <typename T> XMD_HD inline v3<T> cross(v3<T> a, v3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> XMD_HD inline v3<T> unit(v3<T> a) { return norm_inv(a) * a; }
template <typename T> XMD_HD inline T clampT(T v, T lo, T hi) { return v < lo ? lo : (v < hi ? v : hi); }
template <typename T> XMD_HD inline T ipow6(T s) { return s * (s * (s * (s * (s * s)))); }  // math.h:15-19
template <typename T> XMD_HD inline T acos_c(T x) {  // xmd/include/xmd/utils/math.h:41-45
    T x_ = x < T(-1.0) ? T(-1.0) : (x > T(1.0) ? T(1.0) : x);
    return acos(x_);
}
template <typename T, typename U> XMD_HD inline v3<T> cvt(v3<U> a) { return {(T)a.x, (T)a.y, (T)a.z}; }

// FP32 atan2: atan(a) = a P(a^2) on a = min/max of |x|, |y| in [0, 1]
// (degree-8 polynomial in a^2 fitted by iteratively reweighted least
// squares on Chebyshev nodes; max error 6e-9 in exact arithmetic, about
// 1 ulp evaluated in FP32), then the octant reflections.
XMD_HD inline float atan2_f(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 0.0f ? mn * rcp_f(mx) : 0.0f, z = a * a;
    float p = 0.0024567232467234135f;
    p = p * z - 0.014401356689631939f;
    p = p * z + 0.03978123143315315f;
    p = p * z - 0.07234858721494675f;
    p = p * z + 0.10498947650194168f;
    p = p * z - 0.14161230623722076f;
    p = p * z + 0.19985906779766083f;
    p = p * z - 0.33332598209381104f;
    p = p * z + 0.9999998807907104f;
    float r = p * a;
    if (ay > ax) r = 1.5707963267948966f - r;
    if (x < 0.0f) r = 3.141592653589793f - r;
    return copysignf(r, y);
}
// sin and cos of h, |h| <= pi/2: odd / even polynomials (degrees 9 and 10)
// fitted here by iteratively reweighted least squares on Chebyshev nodes;
// error in exact arithmetic 6e-9 (relative, sin) and 3e-10 (absolute, cos),
// about 1 ulp evaluated in FP32.
XMD_HD inline void sincos_half_f(float h, float &sh, float &ch) {
    const float z = h * h;
    float ps = 2.6057807644974673e-06f;
    ps = ps * z - 0.00019809603691101074f;
    ps = ps * z + 0.00833306647837162f;
    ps = ps * z - 0.16666659712791443f;
    sh = h + (h * z) * ps;
    float pc = -2.607710882784886e-07f;
    pc = pc * z + 2.4761886379565112e-05f;
    pc = pc * z - 0.00138884037733078f;
    pc = pc * z + 0.04166664183139801f;
    pc = pc * z - 0.5f;
    ch = 1.0f + z * pc;
}

// Orthorhombic box, xmd/include/xmd/model/box.h:36-49 (cell_inv 0 = open).
struct Box {
    double cx, cy, cz, ix, iy, iz;
};
XMD_HD inline v3<double> wrap(const Box &b, v3<double> v) {
    v3<double> o = v;
    o.x -= b.cx * rint(v.x * b.ix);
    o.y -= b.cy * rint(v.y * b.iy);
    o.z -= b.cz * rint(v.z * b.iz);
    return o;
}
XMD_HD inline v3<double> r_uv(const Box &b, v3<double> u, v3<double> v) { return wrap(b, v - u); }

// --------------------------------------------------------- primitives ---
template <typename T> struct VdV {
    T V, dV;
};
// `d12` = T(12) * depth, the first product of dV, passed in precomputed
// where depth is a kernel constant (the same rounding as forming it here).
template <typename T> XMD_HD inline VdV<T> lj(T depth, T d12, T r_min, T r_inv) {  // lj.h:22-29
    T s = r_inv * r_min, s6 = ipow6(s), s12 = s6 * s6;
    T V = depth * (s12 - T(2.0) * s6);
    T dV = d12 * r_inv * (s6 - s12);
    return {clampT(V, T(-1.0e3), T(1.0e3)), clampT(dV, T(-1.0e3), T(1.0e3))};
}
template <typename T> XMD_HD inline VdV<T> lj(T depth, T r_min, T r_inv) { return lj(depth, T(12.0) * depth, r_min, r_inv); }
template <typename T> XMD_HD inline VdV<T> shifted_lj(T depth, T d12, T r_min, T r_inv) {  // shifted_lj.h:14-21
    T s = r_inv * r_min, s6 = ipow6(s), s12 = s6 * s6;
    T V = depth * (s12 - T(2.0) * s6 + T(1.0));
    T dV = d12 * r_inv * (s6 - s12);
    return {clampT(V, T(-1.0e3), T(1.0e3)), clampT(dV, T(-1.0e3), T(1.0e3))};
}
template <typename T> XMD_HD inline VdV<T> sink_lj(T depth, T r_min, T r_max, T r, T r_inv) {  // sink_lj.h:24-33
    T r_eff = (r < r_min ? r_min : (r < r_max ? r : r_max));
    T s = r_inv * r_eff, s6 = ipow6(s), s12 = s6 * s6;
    T V = depth * (s12 - T(2.0f) * s6);
    T dV = T(12.0f) * depth * r_inv * (s6 - s12);
    dV = (r_min < r && r < r_max) ? T(0) : dV;
    return {clampT(V, T(-1.0e3), T(1.0e3)), clampT(dV, T(-1.0e3), T(1.0e3))};
}
template <typename T> XMD_HD inline VdV<T> harmonic(T H1, T H2, T nat_r, T r) {  // harmonic.h:18-25
    T dr = r - nat_r, dr2 = dr * dr;
    T V = T(0.5) * dr2 * (H1 + H2 * dr2);
    T dV = dr * (H1 + T(2.0) * H2 * dr2);
    return {clampT(V, T(-1.0e3), T(1.0e3)), clampT(dV, T(-1.0e3), T(1.0e3))};
}
// lambda.h:30-56 (+D19)
template <typename T> XMD_HD inline bool lam_supp(T psi0, T alpha, T psi) {
    return fabs(alpha * (psi - psi0)) < T(M_PI);
}
// `api` = alpha / T(pi) (used by the algebraic form), precomputed.
template <typename T> XMD_HD inline VdV<T> lam_eval(T psi0, T alpha, T api, int version, T psi) {
    T s = alpha * (psi - psi0);
    if (version == 0) {
        if constexpr (std::is_same<T, float>::value) {
            // the same function through the half angle (|s| < pi on the
            // support): (1 + cos s)/2 = cos^2(s/2), sin s = 2 sin(s/2) cos(s/2)
            float sh, ch;
            sincos_half_f(0.5f * s, sh, ch);
            return {ch * ch, -alpha * sh * ch};
        } else {
            T val = T(0.5f) * cos(s) + T(0.5f);
            T deriv = T(-0.5f) * alpha * sin(s);
            return {val, deriv};
        }
    }
    T t = fabs(s / T(M_PI));
    T x_inv = rcpT(T(2.0f) * t * t - T(2.0f) * t - T(1));
    T val = (t * t - T(2.0f) * t + T(1.0f)) * x_inv;
    T deriv = T(2.0f) * (t - T(1.0f)) * (t - T(2.0f)) * x_inv * x_inv * api;
    deriv *= (s < T(0.0f) ? T(-1.0f) : T(1.0f));
    return {val, deriv};
}
template <typename T> XMD_HD inline VdV<T> lam_eval(T psi0, T alpha, int version, T psi) {
    return lam_eval(psi0, alpha, alpha / T(M_PI), version, psi);
}

// ------------------------------------------------------------- params ---
// Scalar parameters in the kernels' precision T (narrowed from double on
// the host, like `real` fields set in xmd/src/freeform/main.cpp:345-907).
template <typename T> struct TermParams {
    T H1, H2, e_chi, ang_k, CDA, CDB, CDH;
    T pauli_r_excl, pauli_depth, go_depth;
    // Debye-Hueckel: es_* with es_relative picking the variant (0 const,
    // 1 relative); the engine keeps the constant variant there and the
    // relative one in res_*, so that both run when both are enabled
    T es_Vf, es_sdi;
    int es_relative;
    T res_Vf, res_sdi;
    T bbp_psi0, bbp_alpha, bbp_rmin, bbp_depth;
    T bbm_psi0, bbm_alpha, bbm_rmin, bbm_depth;
    T ss_psi0, ss_alpha;
    int lambda_version;
    // derived: T(12) * depth, alpha / T(pi) (see set_derived)
    T go_d12, pauli_d12, bbp_d12, bbm_d12;
    T bbp_api, bbm_api, ss_api;
    XMD_HD void set_derived() {
        go_d12 = T(12.0) * go_depth;
        pauli_d12 = T(12.0) * pauli_depth;
        bbp_d12 = T(12.0) * bbp_depth;
        bbm_d12 = T(12.0) * bbm_depth;
        bbp_api = bbp_alpha / T(M_PI);
        bbm_api = bbm_alpha / T(M_PI);
        ss_api = ss_alpha / T(M_PI);
    }
};

// ------------------------------------------------------- pair terms ---
// `d` = wrap(r_j - r_g): the displacement from the bead g being summed to
// its partner j.  The returned force is the force on g, which equals the
// reference's F[i1] when g = i1 and F[i2] when g = i2 (d is then -r12).
template <typename T> struct FV {
    v3<T> F;
    T V;
};
// xmd/src/forces/go.cpp:31-46
template <typename T> XMD_HD inline FV<T> go_pair(v3<T> d, T depth, T d12, T nat_dist) {
    T rn = norm_inv(d);
    v3<T> u = d * rn;
    VdV<T> e = lj(depth, d12, nat_dist, rn);
    return {u * e.dV, e.V};
}
// xmd/src/forces/pauli.cpp:32-48
template <typename T> XMD_HD inline FV<T> pauli_pair(v3<T> d, T depth, T d12, T r_excl) {
    T rn = norm_inv(d);
    if (T(1.0f) < rn * r_excl) {
        v3<T> u = d * rn;
        VdV<T> e = shifted_lj(depth, d12, r_excl, rn);
        return {u * e.dV, e.V};
    }
    return {{T(0), T(0), T(0)}, T(0)};
}
// es/const.cpp:13-32 (+D9) and es/relative.cpp:13-32
template <typename T> XMD_HD inline FV<T> es_pair(v3<T> d, T q1q2, T Vf, T sdi, int relative) {
    T r = norm(d), rn = rcpT(r);
    v3<T> u = d * rn;
    T V, dV;
    if (relative) {
        V = Vf * q1q2 * expT(-r * sdi) * rn * rn;
        dV = -V * (sdi + T(2.0f) * rn);
    } else {
        V = Vf * q1q2 * expT(-r * sdi) * rn;
        dV = -V * (sdi + rn);
    }
    return {u * dV, V};
}

// One pseudo-improper dihedral (eval_forces.cpp:27-57) with the gradient
// of psi (D8).  Inputs (FP64): r12 = cen - prev, r23 = next - cen, r13 =
// next - prev, r24 = wrap(partner - cen).  The plane normals (cross
// products of a bond with a vector up to the list radius, which cancel)
// are formed in FP64 in both forms and narrowed; the FP32 form then takes
// psi = atan2 of its sine and cosine parts (well conditioned) and forms
// the gradient in FP32.
template <typename T> struct Psi {
    T psi;
    v3<T> g_prev, g_cen, g_next, g_par;
};
// The partner-independent part of one side: its bonds (FP64), the plane
// normal m = r23 x r12 and the scalars built from them, computed once per
// bead and reused for every partner.
template <typename T> struct PidSide {
    v3<double> r12d, r13d, rmd;
    v3<T> rm, rij, rkj, gi;
    T rm_rn, rkj_n, inv2, pp;
};
template <typename T> XMD_HD inline PidSide<T> pid_side(v3<double> r12d, v3<double> r23d, v3<double> r13d) {
    PidSide<T> S;
    S.r12d = r12d;
    S.r13d = r13d;
    S.rmd = cross(r23d, r12d);
    S.rm = cvt<T>(S.rmd);
    S.rij = cvt<T>(-r23d);
    S.rkj = cvt<T>(-r13d);
    S.rm_rn = norm_inv(S.rm);
    S.rkj_n = norm(S.rkj);
    S.gi = S.rm * (S.rkj_n * S.rm_rn * S.rm_rn);
    S.inv2 = rcpT(S.rkj_n * S.rkj_n);
    S.pp = dot(S.rij, S.rkj) * S.inv2;
    return S;
}
// psi of side S against the partner at r24 = wrap(partner - cen), and its
// gradient.
template <typename T> XMD_HD inline Psi<T> pid_psi(const PidSide<T> &S, v3<double> r24d) {
    const v3<double> r14d = S.r12d + r24d;
    const v3<double> rnd = cross(S.r13d, r14d);
    const v3<T> rn = cvt<T>(rnd);
    const v3<T> rkl = cvt<T>(-r14d);
    const T rn_rn = norm_inv(rn);
    T psi;
    if constexpr (std::is_same<T, double>::value) {
        double cos_psi = dot(S.rm, rn) * S.rm_rn * rn_rn;
        psi = acos_c(cos_psi);
        if (dot(S.rij, rn) < 0.0) psi = -psi;
    } else {
        // sin psi |m||n| = (m.r_kl)|r_kj|: (m x n) = (m.r_kl) r_kj, r_ij.n = m.r_kl;
        // m.r_kl cancels like the cross products, so it is formed in FP64
        psi = atan2_f((float)dot(S.rmd, -r14d) * S.rkj_n, dot(S.rm, rn));
    }
    const v3<T> gl = -(rn * (S.rkj_n * rn_rn * rn_rn));
    const T qq = dot(rkl, S.rkj) * S.inv2;
    const v3<T> gj = (S.pp - T(1.0f)) * S.gi - qq * gl;
    const v3<T> gk = (qq - T(1.0f)) * gl - S.pp * S.gi;
    return {psi, gk, S.gi, gj, gl};
}
// Full PID bundle (eval_forces.cpp:93-140, +D8): sides S1 (bead i1) and S2
// (bead i2), r12 = wrap(r_i2 - r_i1).  out[0..5] = forces on i1p, i1, i1n,
// i2p, i2, i2n.
template <typename T>
XMD_HD inline T pid_bundle(const TermParams<T> &P, T ss_depth, T ss_rmin, T ss_rmax, const PidSide<T> &S1,
                           const PidSide<T> &S2, v3<double> r12d, v3<T> out[6]) {
    Psi<T> g1 = pid_psi<T>(S1, r12d);
    Psi<T> g2 = pid_psi<T>(S2, -r12d);
    T psi1 = g1.psi, psi2 = g2.psi;
    v3<T> r12 = cvt<T>(r12d);
    T r12_n = norm(r12), r12_rn = rcpT(r12_n);
    T A = 0, B = 0, C = 0, V = 0;
    // bb+ then bb-, as in the reference.  A pair is rarely in either and
    // (with the default, disjoint supports) never in both, so the first
    // block evaluates whichever applies with per-lane parameters and the
    // second runs only for a pair in both; the sums keep the bb+, bb-
    // order in every case.
    const bool use_p = lam_supp(P.bbp_psi0, P.bbp_alpha, psi1) && lam_supp(P.bbp_psi0, P.bbp_alpha, psi2);
    const bool use_m = lam_supp(P.bbm_psi0, P.bbm_alpha, psi1) && lam_supp(P.bbm_psi0, P.bbm_alpha, psi2);
    if (use_p || use_m) {
        const T psi0 = use_p ? P.bbp_psi0 : P.bbm_psi0, alpha = use_p ? P.bbp_alpha : P.bbm_alpha;
        const T depth = use_p ? P.bbp_depth : P.bbm_depth, rmin = use_p ? P.bbp_rmin : P.bbm_rmin;
        const T d12 = use_p ? P.bbp_d12 : P.bbm_d12, api = use_p ? P.bbp_api : P.bbm_api;
        VdV<T> l1 = lam_eval(psi0, alpha, api, P.lambda_version, psi1);
        VdV<T> l2 = lam_eval(psi0, alpha, api, P.lambda_version, psi2);
        VdV<T> e = lj(depth, d12, rmin, r12_rn);
        V += l1.V * l2.V * e.V;
        A += l1.dV * l2.V * e.V;
        B += l2.dV * l1.V * e.V;
        C += l1.V * l2.V * e.dV;
    }
    if (use_p && use_m) {
        VdV<T> l1 = lam_eval(P.bbm_psi0, P.bbm_alpha, P.bbm_api, P.lambda_version, psi1);
        VdV<T> l2 = lam_eval(P.bbm_psi0, P.bbm_alpha, P.bbm_api, P.lambda_version, psi2);
        VdV<T> e = lj(P.bbm_depth, P.bbm_d12, P.bbm_rmin, r12_rn);
        V += l1.V * l2.V * e.V;
        A += l1.dV * l2.V * e.V;
        B += l2.dV * l1.V * e.V;
        C += l1.V * l2.V * e.dV;
    }
    if (lam_supp(P.ss_psi0, P.ss_alpha, psi1) && lam_supp(P.ss_psi0, P.ss_alpha, psi2)) {
        VdV<T> l1 = lam_eval(P.ss_psi0, P.ss_alpha, P.ss_api, P.lambda_version, psi1);
        VdV<T> l2 = lam_eval(P.ss_psi0, P.ss_alpha, P.ss_api, P.lambda_version, psi2);
        VdV<T> e = sink_lj(ss_depth, ss_rmin, ss_rmax, r12_n, r12_rn);
        V += l1.V * l2.V * e.V;
        A += l1.dV * l2.V * e.V;
        B += l2.dV * l1.V * e.V;
        C += l1.V * l2.V * e.dV;
    }
    v3<T> u = r12 * r12_rn;
    out[0] = -(A * g1.g_prev);
    out[1] = -(A * g1.g_cen) - B * g2.g_par + C * u;
    out[2] = -(A * g1.g_next);
    out[3] = -(B * g2.g_prev);
    out[4] = -(A * g1.g_par) - B * g2.g_cen - C * u;
    out[5] = -(B * g2.g_next);
    return V;
}

// ---------------------------------------------------- bonded terms ---
// Each returns the forces on all of its beads; a bead kernel keeps its own.
// xmd/src/forces/tether.cpp:21-36 (no wrap)
template <typename T> XMD_HD inline T tether(v3<T> r12, T H1, T H2, T nat, v3<T> out[2]) {
    T r12_n = norm(r12);
    VdV<T> e = harmonic(H1, H2, nat, r12_n);
    v3<T> u = r12 / r12_n;
    out[0] = e.dV * u;
    out[1] = -(e.dV * u);
    return e.V;
}
// xmd/src/forces/chiral.cpp:15-37 (+D16)
template <typename T> XMD_HD inline T chiral(v3<T> r12, v3<T> r23, v3<T> r34, T e_chi, T nat_chir, T nat_factor,
                                            v3<T> out[4]) {
    v3<T> x12_23 = cross(r12, r23), x12_34 = cross(r12, r34), x23_34 = cross(r23, r34);
    T chir = dot(r12, x23_34) * nat_factor;
    T chir_diff = chir - nat_chir;
    T V = T(0.5f) * e_chi * chir_diff * chir_diff;
    T f = e_chi * chir_diff * nat_factor;
    out[0] = f * x23_34;
    out[1] = -(f * (x12_34 + x23_34));
    out[2] = f * (x12_23 + x12_34);
    out[3] = -(f * x12_23);
    return V;
}
// xmd/src/forces/angle/native.cpp:13-39.  In FP32 the normal r12 x r23 is
// formed in FP64 and narrowed (it cancels for a nearly straight angle, and
// the gradient directions are built from it).
template <typename T>
XMD_HD inline T nat_angle(v3<double> r12d, v3<double> r23d, T k, T nat_theta, v3<T> out[3]) {
    const v3<T> r12 = cvt<T>(r12d), r23 = cvt<T>(r23d);
    v3<T> x12_23;
    if constexpr (std::is_same<T, double>::value) x12_23 = cross(r12, r23);
    else x12_23 = cvt<T>(cross(r12d, r23d));
    T r12_rn = norm_inv(r12), r23_rn = norm_inv(r23);
    v3<T> d1 = unit(cross(r12, x12_23)) * r12_rn;
    v3<T> d3 = unit(cross(r23, x12_23)) * r23_rn;
    v3<T> d2 = -d1 - d3;
    T theta;
    if constexpr (std::is_same<T, double>::value) {
        T cos_theta = -dot(r12, r23) * r12_rn * r23_rn;
        theta = acos_c(cos_theta);
    } else {
        theta = atan2_f(norm(x12_23), -dot(r12, r23));  // same angle, FP32-stable
    }
    T dtheta = theta - nat_theta;
    T V = T(0.5) * k * dtheta * dtheta;
    T dV = k * dtheta;
    out[0] = -(dV * d1);
    out[1] = -(dV * d2);
    out[2] = -(dV * d3);
    return V;
}
// The bonds of a dihedral (from FP64 displacements), its two plane normals
// x12_23 = r12 x r23, x23_34 = r23 x r34 and the triple product x12_23.r34.
// In FP32 the normals and the triple product are formed in FP64 and then
// narrowed: for a nearly straight angle the cross product cancels, and the
// dihedral's gradient scales with 1/|x| (FP32 normals put the 100k-bead
// box's dihedral forces 5e-5 off FP64).  The FP64 form is the reference's
// expressions.
template <typename T> struct DihFrame {
    v3<T> r12, r23, r34, x12_23, x23_34;
    T trip;
};
template <typename T> XMD_HD inline DihFrame<T> dih_frame(v3<double> r12d, v3<double> r23d, v3<double> r34d) {
    DihFrame<T> f;
    f.r12 = cvt<T>(r12d);
    f.r23 = cvt<T>(r23d);
    f.r34 = cvt<T>(r34d);
    if constexpr (std::is_same<T, double>::value) {
        f.x12_23 = cross(f.r12, f.r23);
        f.x23_34 = cross(f.r23, f.r34);
        f.trip = dot(f.x12_23, f.r34);
    } else {
        const v3<double> a = cross(r12d, r23d), b = cross(r23d, r34d);
        f.x12_23 = cvt<T>(a);
        f.x23_34 = cvt<T>(b);
        f.trip = (T)dot(a, r34d);
    }
    return f;
}
// complex_native.cpp:13-49 (kind 0) / simple_native.cpp:13-45 (kind 1)
template <typename T>
XMD_HD inline T nat_dihedral(const DihFrame<T> &F, int kind, T CDA, T CDB, T CDH, T nat_phi, v3<T> out[4]) {
    const v3<T> r12 = F.r12, r23 = F.r23, r34 = F.r34, x12_23 = F.x12_23, x23_34 = F.x23_34;
    T x12_23_rn = norm_inv(x12_23), x23_34_rn = norm_inv(x23_34);
    v3<T> x12_23_u = x12_23 * x12_23_rn, x23_34_u = x23_34 * x23_34_rn;
    T phi;
    if constexpr (std::is_same<T, double>::value) {
        T cos_phi = dot(x12_23_u, x23_34_u);
        phi = acos_c(cos_phi);
        if (F.trip < T(0.0f)) phi = -phi;
    } else {
        // same angle, FP32-stable: (x12_23 x x23_34) = (x12_23.r34) r23
        phi = atan2_f(F.trip * norm(r23), dot(x12_23, x23_34));
    }
    T diff = phi - nat_phi;
    T V, dV;
    if (kind == 0) {
        T sin_diff, cos_diff, sin_3diff, cos_3diff;
        if constexpr (std::is_same<T, float>::value) {
            sincosf(diff, &sin_diff, &cos_diff);  // triple angles from the single one
            sin_3diff = sin_diff * (3.0f - 4.0f * sin_diff * sin_diff);
            cos_3diff = cos_diff * (4.0f * cos_diff * cos_diff - 3.0f);
        } else {
            sin_diff = sin(diff), cos_diff = cos(diff);
            sin_3diff = sin(T(3.0f) * diff), cos_3diff = cos(T(3.0f) * diff);
        }
        V = CDA * (T(1.0f) - cos_diff) + CDB * (T(1.0f) - cos_3diff);
        dV = CDA * sin_diff + CDB * T(3.0f) * sin_3diff;
    } else {
        V = T(0.5f) * CDH * diff * diff;
        dV = CDH * diff;
    }
    T r23_n = norm(r23);
    v3<T> d1 = -x12_23_u * r23_n * x12_23_rn;
    v3<T> d4 = x23_34_u * r23_n * x23_34_rn;
    v3<T> df_num = -d1 * dot(r12, r23) + d4 * dot(r23, r34);
    v3<T> df;
    if constexpr (std::is_same<T, float>::value) df = df_num * rcp_f(r23_n * r23_n);
    else df = df_num / (r23_n * r23_n);
    v3<T> d2 = -d1 + df;
    v3<T> d3 = -d4 - df;
    out[0] = -(dV * d1);
    out[1] = -(dV * d2);
    out[2] = -(dV * d3);
    out[3] = -(dV * d4);
    return V;
}

// Heuristic angle (angle/heurestic.cpp:29-59): V = sum_d a_d theta^d,
// d = 0..6, with dV/dtheta by the same Horner sweep.  FP64 in both forms:
// the coefficients alternate in sign up to ~1e5 eps/rad^d while V is a few
// eps, so the sum cancels five digits.
XMD_HD inline double heur_angle(v3<double> r12, v3<double> r23, const double a[7], v3<double> out[3]) {
    v3<double> x12_23 = cross(r12, r23);
    double r12_rn = norm_inv(r12), r23_rn = norm_inv(r23);
    v3<double> d1 = unit(cross(r12, x12_23)) * r12_rn;
    v3<double> d3 = unit(cross(r23, x12_23)) * r23_rn;
    v3<double> d2 = -d1 - d3;
    double theta = acos_c(-dot(r12, r23) * r12_rn * r23_rn);
    double V = 0.0, dV = 0.0;
    for (int d = 6; d >= 0; --d) {
        if (d > 0) dV = (double)d * a[d] + theta * dV;
        V = a[d] + theta * V;
    }
    out[0] = -(dV * d1);
    out[1] = -(dV * d2);
    out[2] = -(dV * d3);
    return V;
}
// Heuristic dihedral (dihedral/heurestic.cpp:29-70): V = c0 + c1 sin phi +
// c2 cos phi + c3 sin^2 + c4 cos^2 + c5 sin cos, phi = +-acos(n1.n2) and
// F = -dV/dphi dphi/dr with the exact derivative (the reference writes
// phi = cos(n1.n2) and 2(c3 + c4) sin cos in dV/dphi; decision D30).
template <typename T>
XMD_HD inline T heur_dihedral(const DihFrame<T> &F, const T c[6], v3<T> out[4]) {
    const v3<T> r12 = F.r12, r23 = F.r23, r34 = F.r34, x12_23 = F.x12_23, x23_34 = F.x23_34;
    T x12_23_rn = norm_inv(x12_23), x23_34_rn = norm_inv(x23_34);
    v3<T> x12_23_u = x12_23 * x12_23_rn, x23_34_u = x23_34 * x23_34_rn;
    T cos_phi = dot(x12_23_u, x23_34_u), sin_phi;
    T r23_n = norm(r23);
    if constexpr (std::is_same<T, double>::value) {
        T phi = acos_c(cos_phi);
        if (F.trip < T(0.0f)) phi = -phi;
        sin_phi = sin(phi);
    } else {
        // sin phi = (x12_23 . r34) |r23| / (|x12_23| |x23_34|): signed and
        // FP32-stable where cos phi is near +-1
        sin_phi = F.trip * r23_n * x12_23_rn * x23_34_rn;
    }
    T sin2 = sin_phi * sin_phi, cos2 = cos_phi * cos_phi, sc = sin_phi * cos_phi;
    T V = c[0] + c[1] * sin_phi + c[2] * cos_phi + c[3] * sin2 + c[4] * cos2 + c[5] * sc;
    T dV = c[1] * cos_phi - c[2] * sin_phi + T(2.0f) * (c[3] - c[4]) * sc + c[5] * (cos2 - sin2);
    v3<T> d1 = -x12_23_u * r23_n * x12_23_rn;
    v3<T> d4 = x23_34_u * r23_n * x23_34_rn;
    v3<T> df_num = -d1 * dot(r12, r23) + d4 * dot(r23, r34);
    v3<T> df;
    if constexpr (std::is_same<T, float>::value) df = df_num * rcp_f(r23_n * r23_n);
    else df = df_num / (r23_n * r23_n);
    v3<T> d2 = -d1 + df;
    v3<T> d3 = -d4 - df;
    out[0] = -(dV * d1);
    out[1] = -(dV * d2);
    out[2] = -(dV * d3);
    out[3] = -(dV * d4);
    return V;
}
// Pair type of the heuristic tables (heur_ang_type, heur_dih_type):
// 0 GLY, 1 PRO, 2 any other residue; type = 3 c1 + c2.  aa_code values of
// GLY and PRO (xmd/include/xmd/types/amino_acid.h, alphabetical codes).
constexpr int AA_GLY = 7, AA_PRO = 14;
XMD_HD inline int heur_class(int aa) { return aa == AA_GLY ? 0 : (aa == AA_PRO ? 1 : 2); }

// ----------------------------------------------------------------- QA ---
// n and h of a bead with prev and next (qa/precompute_nh.cpp:12-20); zero
// otherwise (D17).
XMD_HD inline void qa_nh(bool ok, v3<double> rp, v3<double> rc, v3<double> rx, v3<double> &n, v3<double> &h) {
    if (!ok) {
        n = {0, 0, 0};
        h = {0, 0, 0};
        return;
    }
    v3<double> v1 = rc - rp, v2 = rx - rc;
    n = unit(v2 - v1);
    h = unit(cross(v2, v1));
}
struct QaParams {
    double min_cos_hr, min_cos_hh, max_cos_nr;
    double cycle_time, cycle_time_inv, factor;
    double bb_depth, bb_rmin, bs_depth, bs_rmin;
    double req_max;  // largest required formation distance over contact types
};
// sync_data packed as 4 int8 in a uint32 (back, side_all, side_polar,
// side_hydrophobic), xmd/include/xmd/forces/qa/sync_data.h:8-41.
XMD_HD inline int8_t sync_get(uint32_t s, int k) { return (int8_t)((s >> (8 * k)) & 0xff); }
XMD_HD inline uint32_t sync_sub(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) r |= (uint32_t)(uint8_t)(int8_t)(sync_get(a, k) - sync_get(b, k)) << (8 * k);
    return r;
}
XMD_HD inline uint32_t sync_add(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) r |= (uint32_t)(uint8_t)(int8_t)(sync_get(a, k) + sync_get(b, k)) << (8 * k);
    return r;
}
XMD_HD inline bool sync_valid(uint32_t s) {
    return sync_get(s, 0) >= 0 && sync_get(s, 1) >= 0 && sync_get(s, 2) >= 0 && sync_get(s, 3) >= 0;
}
XMD_HD inline uint32_t sync_pack(int a, int b, int c, int d) {
    return (uint32_t)(uint8_t)a | ((uint32_t)(uint8_t)b << 8) | ((uint32_t)(uint8_t)c << 16) | ((uint32_t)(uint8_t)d << 24);
}
// Candidate test of one free pair, qa/sift_candidates.cpp:15-77.  Returns
// the contact type (0 = not a candidate) and the two sync diffs.
XMD_HD inline int qa_sift(const QaParams &Q, const double *req_min_dist, const int8_t *ptype, v3<double> r12,
                          v3<double> n1, v3<double> h1, v3<double> n2, v3<double> h2, int a1, int a2, uint32_t sync1,
                          uint32_t sync2, uint32_t &d1, uint32_t &d2) {
    double r12_rn = norm_inv(r12);
    v3<double> r12_u = r12 * r12_rn;
    bool back1 = (fabs(dot(h1, r12_u)) > Q.min_cos_hr);
    bool back2 = (fabs(dot(h2, r12_u)) > Q.min_cos_hr);
    bool bb_extra = (fabs(dot(h1, h2)) > Q.min_cos_hh);
    bool side1 = (dot(n1, r12_u) < Q.max_cos_nr);
    bool side2 = (-dot(n2, r12_u) < Q.max_cos_nr);
    int type = 0;
    if (back1 && back2 && bb_extra) type = 1;
    else if (back1 && side2) type = 2;
    else if (side1 && back2) type = 3;
    else if (side1 && side2) type = 4 + a1 + a2 * 20;
    if (type == 0) return 0;
    if (1.0 > r12_rn * req_min_dist[type]) return 0;
    int pt1 = ptype[a1], pt2 = ptype[a2];
    d1 = sync_pack(back1, side1, side1 && pt2 == 1, side1 && pt2 == 2);
    if (!sync_valid(sync_sub(sync1, d1))) return 0;
    d2 = sync_pack(back2, side2, side2 && pt1 == 1, side2 && pt1 == 2);
    if (!sync_valid(sync_sub(sync2, d2))) return 0;
    return type;
}
// LJ of a contact type (D5): returns depth, r_min, r_max of the sink_lj.
XMD_HD inline void qa_lj(const QaParams &Q, const double *ss_depth, const double *ss_rmin, const double *ss_rmax,
                         int type, double &depth, double &r_min, double &r_max) {
    if (type == 1) { depth = Q.bb_depth; r_min = r_max = Q.bb_rmin; return; }
    if (type == 2 || type == 3) { depth = Q.bs_depth; r_min = r_max = Q.bs_rmin; return; }
    int k = type - 4, a1 = k % 20, a2 = k / 20, idx = 20 * a1 + a2;
    depth = ss_depth[idx];
    r_min = ss_rmin[idx];
    r_max = ss_rmax[idx];
}
// Force and transition of one contact, qa/process_contacts.cpp:25-56.
// Returns the force on i1; *next_status: 0/1 status after the step, 2 =
// remove.
XMD_HD inline FV<double> qa_contact(const QaParams &Q, double depth, double r_min, double r_max, v3<double> r12,
                                    double t, double ref_time, int status, int &next_status) {
    double r12_rn = norm_inv(r12);
    v3<double> r12_u = r12 * r12_rn;
    double tt = t - ref_time;
    double saturation = (tt < Q.cycle_time ? tt : Q.cycle_time) * Q.cycle_time_inv;
    if (status == 1) saturation = 1.0f - saturation;
    double r12_n = r12_rn * norm_sq(r12);
    VdV<double> e = sink_lj(depth, r_min, r_max, r12_n, r12_rn);
    FV<double> out;
    out.V = saturation * e.V;
    out.F = saturation * e.dV * r12_u;
    next_status = status;
    if (status == 0 && saturation == 1.0f) {
        if (Q.factor * r_min * r12_rn < 1.0f) next_status = 1;
    } else if (status == 1 && saturation == 0.0f) {
        next_status = 2;
    }
    return out;
}

// ---------------------------------------------------------------- RNG ---
// rand_gen, xmd/include/xmd/random/rand_gen.h:22-35
XMD_HD inline uint64_t rng_out(uint64_t s) { return s * 0xd989bcacc137dcd5ull; }
XMD_HD inline uint64_t rng_step(uint64_t s) {
    s ^= s >> 11;
    s ^= s << 31;
    s ^= s >> 18;
    return s;
}
// normal_x2 / normal of rand_gen (rand_gen.h:42-58) with the corrected
// float libm (bm_math.hpp).  The correction table holds, for each draw value
// k, glibc's corrections of log, sin and cos of that draw (bits 0-1 log,
// 2-3 sin, 4-5 cos; 0 keep, 1 one ulp up, 2 one ulp down).
//
// The float f = (float)y of the engine's double evaluation y can differ from
// glibc's only where y lies near the rounding boundary between f and its
// neighbour on y's side.  a = (half their spacing) - |y - f| measures that
// distance; thr[3 b + j] (j = 0 log, 1 sin, 2 cos) is the largest a /
// spacing among the draw values of bucket b = k >> FIX_BUCKET_SHIFT where
// glibc's function j differs from f, rounded up to float, found on the host
// over every draw value with this same function.  So a draw whose margins
// are all above its bucket's thresholds needs no correction, and only the
// others read the table (about 2 % of log draws and 6 % of angle draws).
constexpr int FIX_BUCKET_SHIFT = 16;
struct NoiseFix {
    const uint8_t *tab;  // nullptr: no correction (correctly rounded noise)
    const float *thr;
};
// a and the spacing (a power of two, so a <= T * spacing is exact); a
// non-finite f gives a = -1, spacing 1 (always below a threshold >= -1).
XMD_HD inline void boundary_margin(double y, float f, double &a, double &sp) {
    if (!(f - f == 0.0f)) {
        a = -1.0;
        sp = 1.0;
        return;
    }
    const double df = (double)f;
    const float g = y > df ? xmd_bm::next_up(f) : xmd_bm::next_down(f);
    sp = fabs((double)g - df);
    a = 0.5 * sp - fabs(y - df);
}
XMD_HD inline bool near_boundary(double y, float f, float T) {
    double a, sp;
    boundary_margin(y, f, a, sp);
    return a <= (double)T * sp;
}
// log of a radius draw (rand_gen.h:47) and sin, cos of an angle draw
// (rand_gen.h:48-49) in double, before rounding to float; v = (float) of the
// draw's upper 32 bits (up to 2^32, which the rounding reaches)
XMD_HD inline double bm_log_d(float v) {
    const float r1 = v * (1.0f / 4294967296.0f);
    return r1 == 0.0f ? -(double)__builtin_inff() : xmd_bm::log_d((double)r1);
}
XMD_HD inline void bm_sincos_d(float v, double &sd, double &cd) {
    const float r2 = v * (1.0f / 4294967296.0f);
    const float t = (float)(2.0 * M_PI) * r2;
    xmd_bm::sincos_red((double)t, sd, cd);
}
XMD_HD inline float bm_radius(float lg) {
    float arg = -2.0f * lg;
    float r = (float)sqrt((double)arg);
    if (r != r) r = 0.0f;
    return r;
}
// The three noise components of one bead from its first state: normal_x2
// from draws 0 (radius) and 1 (angle), normal from draws 2 and 3 (its cosine
// only).
XMD_HD inline void bead_noise(uint64_t s, NoiseFix fx, float out[3]) {
    float u[4];
    for (int k = 0; k < 4; ++k) {
        u[k] = (float)(uint32_t)(rng_out(s) >> 32);
        s = rng_step(s);
    }
    const double l0 = bm_log_d(u[0]), l2 = bm_log_d(u[2]);
    double s1, c1, s3, c3;
    bm_sincos_d(u[1], s1, c1);
    bm_sincos_d(u[3], s3, c3);
    float L0 = (float)l0, L2 = (float)l2, S1 = (float)s1, C1 = (float)c1, C3 = (float)c3;
    if (fx.tab) {
        uint32_t i[4];
        const float *t[4];
        for (int k = 0; k < 4; ++k) {
            i[k] = xmd_bm::draw_index(u[k]);
            t[k] = fx.thr + 3 * (i[k] >> FIX_BUCKET_SHIFT);
        }
        const uint32_t k0 = near_boundary(l0, L0, t[0][0]) ? fx.tab[i[0]] : 0u;
        const uint32_t k1 = near_boundary(s1, S1, t[1][1]) || near_boundary(c1, C1, t[1][2]) ? fx.tab[i[1]] : 0u;
        const uint32_t k2 = near_boundary(l2, L2, t[2][0]) ? fx.tab[i[2]] : 0u;
        const uint32_t k3 = near_boundary(c3, C3, t[3][2]) ? fx.tab[i[3]] : 0u;
        L0 = xmd_bm::apply_fix(L0, k0 & 3u);
        S1 = xmd_bm::apply_fix(S1, (k1 >> 2) & 3u);
        C1 = xmd_bm::apply_fix(C1, (k1 >> 4) & 3u);
        L2 = xmd_bm::apply_fix(L2, k2 & 3u);
        C3 = xmd_bm::apply_fix(C3, (k3 >> 4) & 3u);
    }
    const float r = bm_radius(L0);
    out[0] = r * C1;
    out[1] = r * S1;
    out[2] = bm_radius(L2) * C3;
}

}  // namespace xmd_dev
