// xmd_oracle.cpp -- CPU oracle (test infrastructure, never the product).
//
// A C++ restatement of xmd's per-step force-and-integrate loop, templated
// on the force precision R (double = the FP64 oracle; float = the
// reference's as-shipped `real`, xmd/include/xmd/types/scalar.h:4-5).  The
// Gear derivatives stay double as in xmd/include/xmd/dynamics/lang_pc.h:16.
// All paths below are relative to the reference repository root.
//
// Floating-point expressions are written in the reference's source order
// and the file is compiled with -ffp-contract=off, so every value is the
// ISO-C++ evaluation of the reference expression (GCC's default contraction
// under -march=native would otherwise make results depend on the compiler).
//
// Where the reference cannot be run literally, the oracle makes one of the
// decisions listed below; each is applied identically by the HIP engine.
//   D1  sorted merge against native lists advances its index
//       (xmd/src/nl/legacy_update.cpp:33-55 never advances and hangs).
//   D2  non-native pairs go to non_native (legacy_update.cpp:57-59 pushes
//       them into native_contacts); Pauli, PID and QA filter non_native,
//       Debye-Hueckel filters every nl pair (the filters read a
//       nonexistent nl->particle_pairs, pauli.cpp:15, es_pair.cpp:11,
//       bundle.cpp:11, free_pair.cpp:12).
//   D3  lang_pc's dt, r and num_particles are never assigned and its
//       mass arrays point at temporaries (main.cpp:208-253,217-225): dt
//       from integrator.dt, r = y0, N = the bead count, every mass 1.
//   D4  QA runs precompute_nh -> sift -> process_candidates ->
//       process_contacts every step (main.cpp:1040-1042 runs only two).
//   D5  QA contact LJ: BB -> bb, BS -> bs, SB -> sb, SS(a1,a2) ->
//       ss[20*a1+a2] (lj_variants.h:16-23 is off by one against
//       contact_type.h:10-31).
//   D6  update_free_pairs skips pairs currently in contact
//       (free_pair.cpp:7-20 re-adds them).
//   D7  one sync entry per bead (main.cpp:662-669 appends N more).
//   D8  PID force = -grad of the reference's PID potential: the psi
//       derivatives of eval_forces.cpp:44-53 are scaled by 1/|m| instead
//       of 1/|m|^2, are assigned to the wrong beads, the psi2 block writes
//       dpsi1_* (:83-86) and the radial sign is flipped (:136,139).
//   D9  constant-permittivity DH: V = Vf q1 q2 exp(-r/s)/r with
//       dV/dr = -V(1/s + 1/r) (es/const.cpp:23-24 has 1/r^2 in V).
//   D10 r_excl, atype and max_formation_min_dist set from params.
//   D11 candidates in ascending free-pair index, after precompute_nh.
//   D12 a zero uniform draw gives +-inf noise, reproduced.
//   D13 noise amplitude sqrt(2 kB T gamma / m) without dt^-1/2, as written
//       (lang_pc.cpp:7,13,19).
//   D14 noise is float Box-Muller in every precision mode.
//   D15 synthetic natives use the evaluators' angle convention,
//       theta = acos(-u12.u23) (angle/native.cpp:29-30); models read from
//       PDB files keep the reference's loader convention (D21).
//   D16 chirality force on bead 1 is f*(r23 x r34) (chiral.cpp:33 uses
//       r12 x r23, which is not -dV/dr1 and breaks sum F = 0).
//   D17 beads without prev/next keep n = h = 0 (never written,
//       precompute_nh.cpp:14).
//   D18 chiral quads store (nat_chir, nat_factor) in field order
//       (main.cpp:393 passes them swapped).
//   D19 algebraic lambda: value as written, derivative = d(value)/d(psi)
//       (lambda.h:47 omits alpha and has the wrong polynomial).
//   D23 current disulfide bonds: pairs within 2 * equilibrium distance
//       (harmonic::cutoff, the value main.cpp:592 adds to max_cutoff) plus
//       the list pad; update_nat_ssbonds::cutoff is never set
//       (native_ss.h:44, main.cpp:587-591).
//   D25 solid walls: V = eps / (9 x^9) as written, force -dV/dx n
//       (walls/solid.cpp:24-25 writes dV/dx = V (-9 x) and adds it with
//       the opposite sign).
//   D26 LJ attractive walls: depth 1 eps (lj_attr.h:31 lj_depth is never
//       set and has no parameter); the pair state changes during the
//       step's force pass, as in lj_attr.cpp:14-44.
//   D27 velocity AFM: the harmonic's rest length is 0 (velocity_afm.h:21
//       default-constructs it; main.cpp sets H1, H2 only); a tip without
//       an origin starts at its bead's initial position (main.cpp:884-886
//       reads an undefined pulled_idx), so no force while the tip sits
//       on its bead (the limit of the harmonic's gradient; the literal
//       expression is 0 * inf).  Neither AFM term adds energy
//       (velocity_afm.cpp:21, force_afm.cpp:13).
//   D24 disulfide pairs of the list (native_ssbonds) belong to neither
//       native_contacts nor non_native, so no pair term but the bond sees
//       them (Debye-Hueckel reads native_contacts + non_native, D2).
//   D31 both Debye-Hueckel variants enabled: main.cpp:606-645 builds one
//       pair list with the constant variant's update (cutoff 2 screening
//       distance) and adds each variant's cutoff to max_cutoff; the
//       evaluators' pair pointer is never bound there, so both are taken
//       to run over that list (both are called, main.cpp:1038-1039).
//   D28 heuristic angles and dihedrals: main.cpp never creates their
//       kernels, so their lists are defined here as the model's angles and
//       dihedrals without a native value (the topology's heur_bonded
//       bits), typed by heur_ang_type(a1, a2) of the angle's beads 1, 2
//       and heur_dih_type(a2, a3) of the dihedral's beads 2, 3
//       (angle/heurestic.h, dihedral/heurestic.h argument names).
//   D29 coefficient tables are read by position: 7 angle coefficients
//       a0..a6 (POLY_DEG = 6; defaults.yml rows hold 7 numbers under a
//       header naming a0..a5) and 6 dihedral coefficients const, sin, cos,
//       sin2, cos2, sin_cos (the eval's 6 terms; the header names 5).
//   D30 heuristic dihedral: phi = +-acos(n1.n2) (dihedral/heurestic.cpp:
//       42 writes cos) and dV/dphi exact (:55 has 2(c3 + c4) sin cos for
//       d(c3 sin^2 + c4 cos^2)/dphi = 2(c3 - c4) sin cos).
#include "xmd_oracle.h"

// ORACLE_MUTANT plants one known mistake, for test builds only (the
// shipped oracle is built with 0): tests/test_oracle_value_pins.py builds
// each mutant in a temporary directory and checks that the closed-form
// value pins reject it.  Every mutant keeps F = -grad V, so only a pin of
// V itself can catch it.
//   1  PID: bb+ lambda paired with the bb- LJ and vice versa
//   2  PID: side-side table index transposed (20 a2 + a1)
//   3  PID: sign of psi flipped (with its gradient)
//   4  native angle: the supplement pi - theta (with its gradient)
//   5  native dihedrals: sign of phi flipped (with its gradient)
//   6  chirality: sign of the triple product flipped (with its gradient)
//   7  relative Debye-Hueckel: e^(-r/s)/r instead of e^(-r/s)/r^2
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <set>
#include <string>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

int fail(const std::string &m) {
    g_err = m;
    return XMD_ERR_ARG;
}

// ---------------------------------------------------------------- vec3 ---
// Component expressions follow xmd/include/xmd/types/vec3.h:133-159,327-359.
template <typename R> struct V3 {
    R x, y, z;
};
template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename R> inline V3<R> &operator+=(V3<R> &a, V3<R> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <typename R> inline V3<R> &operator-=(V3<R> &a, V3<R> b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
template <typename R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> inline R norm_sq(V3<R> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <typename R> inline R norm(V3<R> a) { return std::sqrt(norm_sq(a)); }
template <typename R> inline R norm_inv(V3<R> a) { return R(1.0) / norm(a); }
template <typename R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> inline V3<R> unit(V3<R> a) { return norm_inv(a) * a; }
template <typename R> inline R l1_norm(V3<R> a) { return std::abs(a.x) + std::abs(a.y) + std::abs(a.z); }
template <typename R> inline R clampR(R v, R lo, R hi) { return v < lo ? lo : (v < hi ? v : hi); }
template <typename R> inline R acos_c(R x) {  // xmd/include/xmd/utils/math.h:41-45
    R x_ = x < R(-1.0) ? R(-1.0) : (x > R(1.0) ? R(1.0) : x);
    return std::acos(x_);
}
template <typename R> inline R ipow6(R s) { return s * (s * (s * (s * (s * s)))); }  // math.h:15-19

// ---------------------------------------------------------------- box ---
// xmd/include/xmd/model/box.h:12-49
template <typename R> struct Box {
    V3<R> cell{0, 0, 0}, cell_inv{0, 0, 0};
    void set_cell(V3<R> c) {
        cell = c;
        cell_inv = {0, 0, 0};
        if (cell.x != 0.0) cell_inv.x = R(1.0) / cell.x;
        if (cell.y != 0.0) cell_inv.y = R(1.0) / cell.y;
        if (cell.z != 0.0) cell_inv.z = R(1.0) / cell.z;
    }
    V3<R> wrap(V3<R> v) const {
        V3<R> o = v;
        o.x -= cell.x * std::nearbyint(v.x * cell_inv.x);
        o.y -= cell.y * std::nearbyint(v.y * cell_inv.y);
        o.z -= cell.z * std::nearbyint(v.z * cell_inv.z);
        return o;
    }
    V3<R> r_uv(V3<R> u, V3<R> v) const { return wrap(v - u); }
};

// --------------------------------------------------------- primitives ---
template <typename R> struct LJ {  // xmd/include/xmd/forces/primitives/lj.h:22-29
    R depth, r_min;
    std::pair<R, R> operator()(R r_inv) const {
        R s = r_inv * r_min, s6 = ipow6(s), s12 = s6 * s6;
        R V = depth * (s12 - R(2.0) * s6);
        R dV_dr = R(12.0) * depth * r_inv * (s6 - s12);
        V = clampR<R>(V, R(-1.0e3), R(1.0e3));
        dV_dr = clampR<R>(dV_dr, R(-1.0e3), R(1.0e3));
        return {V, dV_dr};
    }
};
template <typename R> struct ShiftedLJ {  // shifted_lj.h:14-21
    R depth, r_min;
    std::pair<R, R> operator()(R r_inv) const {
        R s = r_inv * r_min, s6 = ipow6(s), s12 = s6 * s6;
        R V = depth * (s12 - R(2.0) * s6 + R(1.0));
        R dV_dr = R(12.0) * depth * r_inv * (s6 - s12);
        V = clampR<R>(V, R(-1.0e3), R(1.0e3));
        dV_dr = clampR<R>(dV_dr, R(-1.0e3), R(1.0e3));
        return {V, dV_dr};
    }
};
template <typename R> struct SinkLJ {  // sink_lj.h:24-33
    R depth, r_min, r_max;
    std::pair<R, R> operator()(R r, R r_inv) const {
        R r_eff = (r < r_min ? r_min : (r < r_max ? r : r_max));
        R s = r_inv * r_eff, s6 = ipow6(s), s12 = s6 * s6;
        R V = depth * (s12 - R(2.0f) * s6);
        R dV_dr = R(12.0f) * depth * r_inv * (s6 - s12);
        dV_dr = (r_min < r && r < r_max) ? R(0) : dV_dr;
        V = clampR<R>(V, R(-1.0e3), R(1.0e3));
        dV_dr = clampR<R>(dV_dr, R(-1.0e3), R(1.0e3));
        return {V, dV_dr};
    }
};
template <typename R> struct Harmonic {  // harmonic.h:18-25
    R H1, H2, nat_r;
    std::pair<R, R> operator()(R r) const {
        R dr = r - nat_r, dr2 = dr * dr;
        R V = R(0.5) * dr2 * (H1 + H2 * dr2);
        R dV_dr = dr * (H1 + R(2.0) * H2 * dr2);
        V = clampR<R>(V, R(-1.0e3), R(1.0e3));
        dV_dr = clampR<R>(dV_dr, R(-1.0e3), R(1.0e3));
        return {V, dV_dr};
    }
};
template <typename R> struct Lambda {  // xmd/include/xmd/forces/pid/lambda.h:30-56 (+D19)
    R psi_0, alpha;
    int version;
    bool supp(R psi) const { return std::abs(alpha * (psi - psi_0)) < R(M_PI); }
    std::pair<R, R> operator()(R psi) const {
        if (version == XMD_LAMBDA_COSINE) {
            R s = alpha * (psi - psi_0);
            R val = R(0.5f) * std::cos(s) + R(0.5f);
            R deriv = R(-0.5f) * alpha * std::sin(s);
            return {val, deriv};
        }
        R s = alpha * (psi - psi_0);
        R t = std::abs(s / R(M_PI));
        R x_inv = R(1.0f) / (R(2.0f) * t * t - R(2.0f) * t - R(1));
        R val = (t * t - R(2.0f) * t + R(1.0f)) * x_inv;
        R deriv = R(2.0f) * (t - R(1.0f)) * (t - R(2.0f)) * x_inv * x_inv * (alpha / R(M_PI));
        deriv *= (s < R(0.0f) ? R(-1.0f) : R(1.0f));
        return {val, deriv};
    }
};

// ---------------------------------------------------------------- rng ---
// xmd/include/xmd/random/rand_gen.h:22-58
struct RandGen {
    uint64_t state;
    uint64_t next() {
        uint64_t res = state * 0xd989bcacc137dcd5ull;
        state ^= state >> 11;
        state ^= state << 31;
        state ^= state >> 18;
        return res;
    }
    float uniformf() {
        float inv = 1.0f / (float)(1ull << 32);
        uint64_t val = next();
        return (float)(val >> 32ull) * inv;
    }
    float normalf() {
        float r1 = uniformf(), r2 = uniformf();
        float r = std::sqrt(-2.0f * std::log(r1));
        if (std::isnan(r)) r = 0.0f;
        float t = (float)(2.0 * M_PI) * r2;
        return r * std::cos(t);
    }
    std::pair<float, float> normal_x2f() {
        float r1 = uniformf(), r2 = uniformf();
        float r = std::sqrt(-2.0f * std::log(r1));
        if (std::isnan(r)) r = 0.0f;
        float t = (float)(2.0 * M_PI) * r2;
        return {r * std::cos(t), r * std::sin(t)};
    }
    // double forms (uniform<double>, uniform(a, b), normal_x2, normal, sphere)
    double uniformd() {
        double inv = 1.0 / (double)(1ull << 32);
        return (double)(next() >> 32ull) * inv;
    }
    double uniformd(double a, double b) { return (b - a) * uniformd() + a; }
    void sphere(double v[3]) {
        double r1 = uniformd(), r2 = uniformd();
        double r = std::sqrt(-2.0 * std::log(r1));
        if (std::isnan(r)) r = 0.0;
        double t = (2.0 * M_PI) * r2;
        v[0] = r * std::cos(t);
        v[1] = r * std::sin(t);
        r1 = uniformd(), r2 = uniformd();
        r = std::sqrt(-2.0 * std::log(r1));
        if (std::isnan(r)) r = 0.0;
        t = (2.0 * M_PI) * r2;
        v[2] = r * std::cos(t);
        double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; ++k) v[k] /= n;
    }
};

// ---------------------------------------------------------- the system ---
struct Sync {
    int8_t v[4];  // back, side_all, side_polar, side_hydrophobic (sync_data.h:8-9)
    bool valid() const { return v[0] >= 0 && v[1] >= 0 && v[2] >= 0 && v[3] >= 0; }
};
inline Sync sub(Sync a, Sync b) {
    Sync r;
    for (int k = 0; k < 4; ++k) r.v[k] = (int8_t)(a.v[k] - b.v[k]);
    return r;
}
inline Sync add(Sync a, Sync b) {
    Sync r;
    for (int k = 0; k < 4; ++k) r.v[k] = (int8_t)(a.v[k] + b.v[k]);
    return r;
}

template <typename R> struct NlPair {
    int i1, i2;
    R orig_dist;
    bool native;  // native_contacts or native_ssbonds (not non_native)
    bool ss;      // native_ssbonds
};
template <typename R> struct NatCont {
    int i1, i2;
    R nat_dist;
};
template <typename R> struct EsPair {
    int i1, i2;
    R q1_x_q2;
};
struct PidBundle {
    int i1, i2;
    int16_t type;
};
struct FreePair {
    int i1, i2;
    bool vacant;
};
struct Cand {
    int i1, i2, fp_idx;
    int16_t type;
    Sync d1, d2;
};
template <typename R> struct Contact {
    int i1, i2;
    int16_t type;
    int8_t status;
    R ref_time;
    Sync d1, d2;
    bool vacant;
};
template <typename R> struct Tether {
    int i1, i2;
    R nat;
};
template <typename R> struct Quad {
    int i1, i2, i3, i4;
    R nat_chir, nat_factor;
};
template <typename R> struct Angle {
    int i1, i2, i3;
    R nat_theta;
};
template <typename R> struct Dih {
    int i1, i2, i3, i4;
    R nat_phi;
};
// heur_ang / heur_dih records (angle/heurestic.h, dihedral/heurestic.h):
// beads and the pair type 3 c1 + c2 (c: 0 GLY, 1 PRO, 2 other).
struct HeurAngle {
    int i1, i2, i3, type;
};
struct HeurDih {
    int i1, i2, i3, i4, type;
};

// Parameters narrowed to `real`, mirroring the wiring in
// xmd/src/freeform/main.cpp:345-907.
template <typename R> struct Params {
    bool lang_on;
    double dt;
    R gamma_factor;
    R pad_factor, max_cutoff;
    bool tether_on, chir_on, ang_on, cnd_on, snd_on, pauli_on, go_on, ces_on, res_on, qa_on, pid_on, ss_on;
    R ss_H1, ss_equil;
    bool solid_on, ljw_on, vafm_on, fafm_on;
    R solid_eps, solid_cut, ljw_min_dist, ljw_factor, ljw_cycle, ljw_cycle_inv, vafm_H1, vafm_H2;
    std::vector<std::array<R, 6>> solid_walls, lj_walls;  // origin, normal
    R H1, H2, def_len, e_chi, ang_k, CDA, CDB, CDH;
    bool hang_on, hdih_on;
    R hang_c[9][7], hdih_c[9][6];  // eval_heurestic_*_forces tables by pair type
    R pauli_r_excl, pauli_depth, go_depth;
    R ces_Vf, ces_sdi, res_Vf, res_sdi, es_cutoff;
    R q[XMD_NUM_AA];
    LJ<R> bb, bs, sb;
    SinkLJ<R> ss[XMD_NUM_SS];
    Lambda<R> bbp_lam, bbm_lam, ss_lam;
    LJ<R> bbp_lj, bbm_lj;
    R pid_cutoff;
    R cycle_time, cycle_time_inv, factor;
    R min_cos_hr, min_cos_hh, max_cos_nr;
    R req_min_dist[XMD_QA_NUM_TYPES];
    R max_formation_min_dist;
    int8_t ptype[XMD_NUM_AA];
    Sync limits[XMD_NUM_AA];
};

template <typename R> SinkLJ<R> qa_lj(const Params<R> &P, int type) {  // D5
    if (type == 1) return {P.bb.depth, P.bb.r_min, P.bb.r_min};
    if (type == 2) return {P.bs.depth, P.bs.r_min, P.bs.r_min};
    if (type == 3) return {P.sb.depth, P.sb.r_min, P.sb.r_min};
    int k = type - 4, a1 = k % XMD_NUM_AA, a2 = k / XMD_NUM_AA;
    return P.ss[XMD_NUM_AA * a1 + a2];
}

template <typename R> struct Replica {
    int n = 0, off = 0;
    std::vector<int8_t> atype;
    std::vector<int> prev, next, chain, seq;
    Box<R> box;
    R temperature;
    RandGen gen{0};
    std::vector<V3<R>> r, v, F;
    std::vector<V3<double>> y[6];
    R t = 0;
    double true_t = 0;
    bool first_time = true, invalid = true;
    // nl_data (xmd/include/xmd/nl/data.h:8-14); `nl` keeps every pair in
    // canonical order with a native flag, i.e. native_contacts + non_native.
    std::vector<NlPair<R>> nl;
    R orig_pad = 0, ref_t = 0;
    std::vector<V3<R>> orig_r;
    Box<R> orig_box;
    std::vector<NatCont<R>> all_nat, go_cur;
    std::vector<std::pair<int, int>> all_nat_nl;  // sorted (i1,i2)
    std::vector<std::pair<int, int>> all_ss, ss_cur;  // sorted (i1,i2); current
    struct LjWall {
        int8_t status = 0;  // 0 FREE, 1 FORMING_OR_FORMED, 2 BREAKING
        V3<R> joint{0, 0, 0};
        R ref_t = 0;
    };
    std::vector<LjWall> ljw;  // bead * n_lj_walls + wall
    struct VelTip {
        int i;
        V3<R> orig, v;
    };
    struct ForceTip {
        int i;
        V3<R> F;
    };
    std::vector<VelTip> vtips;
    std::vector<ForceTip> ftips;
    std::vector<std::pair<int, int>> pauli;
    std::vector<EsPair<R>> es;
    std::vector<PidBundle> pid;
    std::vector<FreePair> fp;
    std::vector<Cand> cand;
    std::vector<Contact<R>> cont;
    std::vector<Sync> sync;
    std::vector<V3<R>> nvec, hvec;
    std::vector<Tether<R>> tethers;
    std::vector<Quad<R>> quads;
    std::vector<Angle<R>> angles;
    std::vector<Dih<R>> dihs;
    std::vector<HeurAngle> hangles;
    std::vector<HeurDih> hdihs;
    R V_last = 0;
    int64_t rebuilds = 0, steps = 0;
};

// --------------------------------------------------------- force terms ---
// Each eval_* accumulates into F and *V exactly like eval_X_forces::iter.
template <typename R> void eval_tether(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.tethers.size()); ++k) {
        auto const &t = s.tethers[k];  // xmd/src/forces/tether.cpp:21-36
        V3<R> r12 = s.r[t.i2] - s.r[t.i1];
        R r12_n = norm(r12);
        auto [V_, dV_dr] = Harmonic<R>{P.H1, P.H2, t.nat}(r12_n);
        V3<R> r12_u = r12 / r12_n;
        *V += V_;
        F[t.i1] += dV_dr * r12_u;
        F[t.i2] -= dV_dr * r12_u;
    }
}
template <typename R> void eval_chiral(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.quads.size()); ++k) {
        auto const &q = s.quads[k];  // xmd/src/forces/chiral.cpp:15-37 (+D16, D18)
        V3<R> r1 = s.r[q.i1], r2 = s.r[q.i2], r3 = s.r[q.i3], r4 = s.r[q.i4];
        V3<R> r12 = r2 - r1, r23 = r3 - r2, r34 = r4 - r3;
        V3<R> x12_23 = cross(r12, r23), x12_34 = cross(r12, r34), x23_34 = cross(r23, r34);
        R chir = dot(r12, x23_34) * q.nat_factor;
        if (ORACLE_MUTANT == 6) chir = -chir;
        R chir_diff = chir - q.nat_chir;
        *V += R(0.5f) * P.e_chi * chir_diff * chir_diff;
        R f = P.e_chi * chir_diff * q.nat_factor;
        if (ORACLE_MUTANT == 6) f = -f;
        F[q.i1] += f * x23_34;
        F[q.i2] -= f * (x12_34 + x23_34);
        F[q.i3] += f * (x12_23 + x12_34);
        F[q.i4] -= f * x12_23;
    }
}
template <typename R> void eval_angle(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.angles.size()); ++k) {
        auto const &a = s.angles[k];  // xmd/src/forces/angle/native.cpp:13-39
        V3<R> r1 = s.r[a.i1], r2 = s.r[a.i2], r3 = s.r[a.i3];
        V3<R> r12 = r2 - r1, r23 = r3 - r2;
        V3<R> x12_23 = cross(r12, r23);
        R r12_rn = norm_inv(r12), r23_rn = norm_inv(r23);
        V3<R> dtheta_dr1 = unit(cross(r12, x12_23)) * r12_rn;
        V3<R> dtheta_dr3 = unit(cross(r23, x12_23)) * r23_rn;
        V3<R> dtheta_dr2 = -dtheta_dr1 - dtheta_dr3;
        R cos_theta = -dot(r12, r23) * r12_rn * r23_rn;
        R theta = acos_c(cos_theta);
        if (ORACLE_MUTANT == 4) theta = R(M_PI) - theta;
        R dtheta = theta - a.nat_theta;
        *V += R(0.5) * P.ang_k * dtheta * dtheta;
        R dV_dtheta = P.ang_k * dtheta;
        if (ORACLE_MUTANT == 4) dV_dtheta = -dV_dtheta;
        F[a.i1] -= dV_dtheta * dtheta_dr1;
        F[a.i2] -= dV_dtheta * dtheta_dr2;
        F[a.i3] -= dV_dtheta * dtheta_dr3;
    }
}
// Shared geometry of the native dihedral terms
// (complex_native.cpp:19-43, simple_native.cpp:17-37).
template <typename R> struct DihGeo {
    R phi;
    V3<R> d1, d2, d3, d4;
};
template <typename R> DihGeo<R> dih_geo(V3<R> r1, V3<R> r2, V3<R> r3, V3<R> r4) {
    V3<R> r12 = r2 - r1, r23 = r3 - r2, r34 = r4 - r3;
    V3<R> x12_23 = cross(r12, r23), x23_34 = cross(r23, r34);
    R x12_23_rn = norm_inv(x12_23), x23_34_rn = norm_inv(x23_34);
    V3<R> x12_23_u = x12_23 * x12_23_rn, x23_34_u = x23_34 * x23_34_rn;
    R cos_phi = dot(x12_23_u, x23_34_u);
    R phi = acos_c(cos_phi);
    if (dot(x12_23, r34) < R(0.0f)) phi = -phi;
    R r23_n = norm(r23);
    V3<R> dphi_dr1 = -x12_23_u * r23_n * x12_23_rn;
    V3<R> dphi_dr4 = x23_34_u * r23_n * x23_34_rn;
    V3<R> df = (-dphi_dr1 * dot(r12, r23) + dphi_dr4 * dot(r23, r34)) / (r23_n * r23_n);
    V3<R> dphi_dr2 = -dphi_dr1 + df;
    V3<R> dphi_dr3 = -dphi_dr4 - df;
    if (ORACLE_MUTANT == 5) return {-phi, -dphi_dr1, -dphi_dr2, -dphi_dr3, -dphi_dr4};
    return {phi, dphi_dr1, dphi_dr2, dphi_dr3, dphi_dr4};
}
template <typename R> void eval_cnd(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.dihs.size()); ++k) {
        auto const &d = s.dihs[k];  // complex_native.cpp:13-49
        auto g = dih_geo(s.r[d.i1], s.r[d.i2], s.r[d.i3], s.r[d.i4]);
        R diff = g.phi - d.nat_phi;
        R sin_diff = std::sin(diff), cos_diff = std::cos(diff);
        R sin_3diff = std::sin(R(3.0f) * diff), cos_3diff = std::cos(R(3.0f) * diff);
        *V += P.CDA * (R(1.0f) - cos_diff) + P.CDB * (R(1.0f) - cos_3diff);
        R dV_dphi = P.CDA * sin_diff + P.CDB * R(3.0f) * sin_3diff;
        F[d.i1] -= dV_dphi * g.d1;
        F[d.i2] -= dV_dphi * g.d2;
        F[d.i3] -= dV_dphi * g.d3;
        F[d.i4] -= dV_dphi * g.d4;
    }
}
template <typename R> void eval_snd(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.dihs.size()); ++k) {
        auto const &d = s.dihs[k];  // simple_native.cpp:13-45
        auto g = dih_geo(s.r[d.i1], s.r[d.i2], s.r[d.i3], s.r[d.i4]);
        R diff = g.phi - d.nat_phi;
        *V += R(0.5f) * P.CDH * diff * diff;
        R dV_dphi = P.CDH * diff;
        F[d.i1] -= dV_dphi * g.d1;
        F[d.i2] -= dV_dphi * g.d2;
        F[d.i3] -= dV_dphi * g.d3;
        F[d.i4] -= dV_dphi * g.d4;
    }
}
// Heuristic angles (xmd/src/forces/angle/heurestic.cpp:29-59): degree-6
// polynomial in theta, value and derivative by one Horner sweep from the
// top coefficient, as the reference loops d = POLY_DEG..0.
template <typename R> void eval_heur_angle(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.hangles.size()); ++k) {
        auto const &a = s.hangles[k];
        V3<R> r1 = s.r[a.i1], r2 = s.r[a.i2], r3 = s.r[a.i3];
        V3<R> r12 = r2 - r1, r23 = r3 - r2;
        V3<R> x12_23 = cross(r12, r23);
        R r12_rn = norm_inv(r12), r23_rn = norm_inv(r23);
        V3<R> dtheta_dr1 = unit(cross(r12, x12_23)) * r12_rn;
        V3<R> dtheta_dr3 = unit(cross(r23, x12_23)) * r23_rn;
        V3<R> dtheta_dr2 = -dtheta_dr1 - dtheta_dr3;
        R cos_theta = -dot(r12, r23) * r12_rn * r23_rn;
        R theta = acos_c(cos_theta);
        R angle_V = 0, dV_dtheta = 0;
        for (int d = 6; d >= 0; --d) {
            R coeff = P.hang_c[a.type][d];
            if (d > 0) dV_dtheta = (R)d * coeff + theta * dV_dtheta;
            angle_V = coeff + theta * angle_V;
        }
        *V += angle_V;
        F[a.i1] -= dV_dtheta * dtheta_dr1;
        F[a.i2] -= dV_dtheta * dtheta_dr2;
        F[a.i3] -= dV_dtheta * dtheta_dr3;
    }
}
// Heuristic dihedrals (xmd/src/forces/dihedral/heurestic.cpp:29-70) with
// D30: phi = +-acos(n1.n2) (the reference writes cos) and the exact
// dV/dphi (the reference's 2(c3 + c4) sin cos term has c4's sign flipped).
template <typename R> void eval_heur_dih(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.hdihs.size()); ++k) {
        auto const &d = s.hdihs[k];
        auto g = dih_geo(s.r[d.i1], s.r[d.i2], s.r[d.i3], s.r[d.i4]);
        V3<R> r12 = s.r[d.i2] - s.r[d.i1], r23 = s.r[d.i3] - s.r[d.i2], r34 = s.r[d.i4] - s.r[d.i3];
        V3<R> x12_23 = cross(r12, r23), x23_34 = cross(r23, r34);
        R cos_phi = dot(x12_23 * norm_inv(x12_23), x23_34 * norm_inv(x23_34));
        R sin_phi = std::sin(g.phi);
        R sin2 = sin_phi * sin_phi, cos2 = cos_phi * cos_phi, sc = sin_phi * cos_phi;
        const R *c = P.hdih_c[d.type];
        *V += c[0] + c[1] * sin_phi + c[2] * cos_phi + c[3] * sin2 + c[4] * cos2 + c[5] * sc;
        R dV_dphi = c[1] * cos_phi - c[2] * sin_phi + R(2.0f) * (c[3] - c[4]) * sc + c[5] * (cos2 - sin2);
        F[d.i1] -= dV_dphi * g.d1;
        F[d.i2] -= dV_dphi * g.d2;
        F[d.i3] -= dV_dphi * g.d3;
        F[d.i4] -= dV_dphi * g.d4;
    }
}
template <typename R> void eval_pauli(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.pauli.size()); ++k) {
        auto const &pr = s.pauli[k];  // xmd/src/forces/pauli.cpp:32-48
        int i1 = pr.first, i2 = pr.second;
        V3<R> r12 = s.box.r_uv(s.r[i1], s.r[i2]);
        R r12_rn = norm_inv(r12);
        if (R(1.0f) < r12_rn * P.pauli_r_excl) {
            V3<R> r12_u = r12 * r12_rn;
            auto [V_, dV_dr] = ShiftedLJ<R>{P.pauli_depth, P.pauli_r_excl}(r12_rn);
            *V += V_;
            F[i1] += r12_u * dV_dr;
            F[i2] -= r12_u * dV_dr;
        }
    }
}
template <typename R> void eval_go(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.go_cur.size()); ++k) {
        auto const &c = s.go_cur[k];  // xmd/src/forces/go.cpp:31-46
        V3<R> r12 = s.box.r_uv(s.r[c.i1], s.r[c.i2]);
        R r12_rn = norm_inv(r12);
        V3<R> r12_u = r12 * r12_rn;
        auto [V_, dV_dr] = LJ<R>{P.go_depth, c.nat_dist}(r12_rn);
        *V += V_;
        F[c.i1] += r12_u * dV_dr;
        F[c.i2] -= r12_u * dV_dr;
    }
}
template <typename R> void eval_nat_ss(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.ss_cur.size()); ++k) {
        int i1 = s.ss_cur[k].first, i2 = s.ss_cur[k].second;  // xmd/src/forces/native_ss.cpp:27-40
        V3<R> r12 = s.box.r_uv(s.r[i1], s.r[i2]);
        R r12_n = norm(r12);
        auto [V_, dV_dr] = Harmonic<R>{P.ss_H1, R(0.0f), P.ss_equil}(r12_n);
        V3<R> r12_u = r12 / r12_n;
        *V += V_;
        F[i1] += dV_dr * r12_u;
        F[i2] -= dV_dr * r12_u;
    }
}
// xmd/src/forces/walls/solid.cpp:19-30 (+D25)
template <typename R> void eval_solid_walls(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0,
                                            size_t hi = SIZE_MAX) {
    for (size_t i = lo; i < std::min(hi, (size_t)s.n); ++i) {
        V3<R> pr = s.box.wrap(s.r[i]);
        for (auto const &w : P.solid_walls) {
            V3<R> o{w[0], w[1], w[2]}, nn{w[3], w[4], w[5]};
            R x = dot(nn, pr - o);
            if (x < P.solid_cut) {
                R x3 = x * x * x;
                R V_ = P.solid_eps / (R(9.0) * (x3 * x3 * x3));
                R dV_dx = -R(9.0) * V_ / x;
                *V += V_;
                F[i] -= dV_dx * nn;
            }
        }
    }
}
// xmd/src/forces/walls/lj_attr.cpp:7-49 (+D26); mutate: the step's pass
template <typename R> void eval_lj_walls(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, bool mutate,
                                         size_t lo = 0, size_t hi = SIZE_MAX) {
    const size_t nw = P.lj_walls.size();
    for (size_t k = lo; k < std::min(hi, s.ljw.size()); ++k) {
        auto &pr = s.ljw[k];
        const int i = (int)(k / nw);
        auto const &w = P.lj_walls[k % nw];
        V3<R> o{w[0], w[1], w[2]}, nn{w[3], w[4], w[5]};
        V3<R> r_ = s.box.wrap(s.r[i]);
        if (pr.status == 0) {
            R x = dot(nn, r_ - o);
            if (std::abs(x) < P.ljw_min_dist && mutate) {
                pr.status = 1;
                pr.joint = r_ - nn * x;
                pr.ref_t = s.t;
            }
            continue;
        }
        R saturation = std::min(s.t - pr.ref_t, P.ljw_cycle) * P.ljw_cycle_inv;
        if (pr.status == 2) saturation = R(1.0) - saturation;
        V3<R> r12 = pr.joint - r_;
        R r12_rn = norm_inv(r12);
        auto [V_, dV_dr] = LJ<R>{R(1.0), P.ljw_min_dist}(r12_rn);
        *V += V_ * saturation;
        V3<R> r12_u = r12 * r12_rn;
        F[i] += saturation * dV_dr * r12_u;
        if (!mutate) continue;
        if (pr.status == 1 && saturation == R(1.0f)) {
            if (P.ljw_factor * P.ljw_min_dist * r12_rn < R(1.0f)) {
                pr.status = 2;
                pr.ref_t = s.t;
            }
        } else if (pr.status == 2 && saturation == R(0.0f)) {
            pr.status = 0;
        }
    }
}
// xmd/src/forces/velocity_afm.cpp:12-24, force_afm.cpp:11-14 (+D27)
template <typename R> void eval_vel_afm(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, size_t lo = 0,
                                        size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.vtips.size()); ++k) {
        auto const &tip = s.vtips[k];
        V3<R> cur = tip.orig + s.t * tip.v;
        V3<R> r_afm = cur - s.r[tip.i];
        R r_afm_n = norm(r_afm);
        if (r_afm_n == R(0.0)) continue;  // D27: the force's limit, 0
        R r_afm_rn = R(1.0f) / r_afm_n;
        V3<R> r_afm_u = r_afm * r_afm_rn;
        auto [V_, dV_dr] = Harmonic<R>{P.vafm_H1, P.vafm_H2, R(0.0)}(r_afm_n);
        (void)V_;
        F[tip.i] += dV_dr * r_afm_u;
    }
}
template <typename R> void eval_force_afm(Replica<R> &s, std::vector<V3<R>> &F, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.ftips.size()); ++k) F[s.ftips[k].i] += s.ftips[k].F;
}
template <typename R> void eval_es(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, bool relative, size_t lo = 0,
             size_t hi = SIZE_MAX) {
    R Vf = relative ? P.res_Vf : P.ces_Vf, sdi = relative ? P.res_sdi : P.ces_sdi;
    for (size_t k = lo; k < std::min(hi, s.es.size()); ++k) {  // es/const.cpp:13-32 (+D9), es/relative.cpp:13-32
        auto const &e = s.es[k];
        V3<R> r12 = s.box.r_uv(s.r[e.i1], s.r[e.i2]);
        R r12_n = norm(r12), r12_rn = R(1.0f) / r12_n;
        V3<R> r12_u = r12 * r12_rn;
        R Vij, dVij_dr;
        if (relative) {
            Vij = Vf * e.q1_x_q2 * std::exp(-r12_n * sdi) * r12_rn * r12_rn;
            dVij_dr = -Vij * (sdi + R(2.0f) * r12_rn);
            if (ORACLE_MUTANT == 7) {
                Vij = Vf * e.q1_x_q2 * std::exp(-r12_n * sdi) * r12_rn;
                dVij_dr = -Vij * (sdi + r12_rn);
            }
        } else {
            Vij = Vf * e.q1_x_q2 * std::exp(-r12_n * sdi) * r12_rn;
            dVij_dr = -Vij * (sdi + r12_rn);
        }
        *V += Vij;
        V3<R> f = r12_u * dVij_dr;
        F[e.i1] += f;
        F[e.i2] -= f;
    }
}

// One pseudo-improper dihedral and its gradient (D8).  Beads: prev, cen,
// next of one chain and the partner `par`; geometry as
// xmd/src/forces/pid/eval_forces.cpp:27-57.
template <typename R> struct PsiGeo {
    R psi;
    V3<R> g_prev, g_cen, g_next, g_par;
};
template <typename R> PsiGeo<R> pid_psi(const Box<R> &box, V3<R> r1_, V3<R> r2_, V3<R> r3_, V3<R> r4_) {
    V3<R> r24 = box.r_uv(r2_, r4_);
    V3<R> r12 = r2_ - r1_;
    V3<R> r23 = r3_ - r2_;
    V3<R> r13 = r3_ - r1_;
    V3<R> r14 = r12 + r24;
    V3<R> rij = -r23, rkj = -r13, rkl = -r14;
    V3<R> rm = cross(r23, r12);
    V3<R> rn = cross(r13, r14);
    R rm_rn = norm_inv(rm), rn_rn = norm_inv(rn);
    R rkj_n = norm(rkj);
    R cos_psi = dot(rm, rn) * rm_rn * rn_rn;
    R psi = acos_c(cos_psi);
    if (dot(rij, rn) < R(0.0f)) psi = -psi;
    // gradient of psi: i = cen (r2_), j = next (r3_), k = prev (r1_), l = par
    V3<R> gi = rm * (rkj_n * rm_rn * rm_rn);
    V3<R> gl = -(rn * (rkj_n * rn_rn * rn_rn));
    R inv2 = R(1) / (rkj_n * rkj_n);
    R p = dot(rij, rkj) * inv2, q = dot(rkl, rkj) * inv2;
    V3<R> gj = (p - R(1)) * gi - q * gl;
    V3<R> gk = (q - R(1)) * gl - p * gi;
    if (ORACLE_MUTANT == 3) return {-psi, -gk, -gi, -gj, -gl};
    return {psi, gk, gi, gj, gl};
}

template <typename R>
void pid_bundle_eval(const Params<R> &P, const Box<R> &box, const std::vector<V3<R>> &r, const std::vector<int> &prev,
                     const std::vector<int> &next, int i1, int i2, int type, R &V_, V3<R> out[6]) {
    // out: forces on i1p, i1, i1n, i2p, i2, i2n
    int i1p = prev[i1], i1n = next[i1], i2p = prev[i2], i2n = next[i2];
    auto g1 = pid_psi(box, r[i1p], r[i1], r[i1n], r[i2]);
    auto g2 = pid_psi(box, r[i2p], r[i2], r[i2n], r[i1]);
    R psi1 = g1.psi, psi2 = g2.psi;
    V3<R> r12 = box.r_uv(r[i1], r[i2]);
    R r12_n = norm(r12), r12_rn = R(1.0f) / r12_n;
    R A = 0, B = 0, C = 0;
    V_ = 0;
    if (P.bbp_lam.supp(psi1) && P.bbp_lam.supp(psi2)) {
        auto [lam1, deriv1] = P.bbp_lam(psi1);
        auto [lam2, deriv2] = P.bbp_lam(psi2);
        auto [lj_V, lj_dV_dr] = P.bbp_lj(r12_rn);
        V_ += lam1 * lam2 * lj_V;
        A += deriv1 * lam2 * lj_V;
        B += deriv2 * lam1 * lj_V;
        C += lam1 * lam2 * lj_dV_dr;
    }
    if (P.bbm_lam.supp(psi1) && P.bbm_lam.supp(psi2)) {
        auto [lam1, deriv1] = P.bbm_lam(psi1);
        auto [lam2, deriv2] = P.bbm_lam(psi2);
        auto [lj_V, lj_dV_dr] = P.bbm_lj(r12_rn);
        V_ += lam1 * lam2 * lj_V;
        A += deriv1 * lam2 * lj_V;
        B += deriv2 * lam1 * lj_V;
        C += lam1 * lam2 * lj_dV_dr;
    }
    SinkLJ<R> ssl = P.ss[type];
    if (P.ss_lam.supp(psi1) && P.ss_lam.supp(psi2)) {
        auto [lam1, deriv1] = P.ss_lam(psi1);
        auto [lam2, deriv2] = P.ss_lam(psi2);
        auto [lj_V, lj_dV_dr] = ssl(r12_n, r12_rn);
        V_ += lam1 * lam2 * lj_V;
        A += deriv1 * lam2 * lj_V;
        B += deriv2 * lam1 * lj_V;
        C += lam1 * lam2 * lj_dV_dr;
    }
    V3<R> u = r12 * r12_rn;
    out[0] = -(A * g1.g_prev);
    out[1] = -(A * g1.g_cen) - B * g2.g_par + C * u;
    out[2] = -(A * g1.g_next);
    out[3] = -(B * g2.g_prev);
    out[4] = -(A * g1.g_par) - B * g2.g_cen - C * u;
    out[5] = -(B * g2.g_next);
}

template <typename R> void eval_pid(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, size_t lo = 0, size_t hi = SIZE_MAX) {
    for (size_t k = lo; k < std::min(hi, s.pid.size()); ++k) {
        auto const &b = s.pid[k];  // xmd/src/forces/pid/eval_forces.cpp:14-141 (+D8)
        R V_;
        V3<R> out[6];
        pid_bundle_eval(P, s.box, s.r, s.prev, s.next, b.i1, b.i2, b.type, V_, out);
        *V += V_;
        F[s.prev[b.i1]] += out[0];
        F[b.i1] += out[1];
        F[s.next[b.i1]] += out[2];
        F[s.prev[b.i2]] += out[3];
        F[b.i2] += out[4];
        F[s.next[b.i2]] += out[5];
    }
}

// ------------------------------------------------------------------ QA ---
template <typename R> void qa_precompute_nh(Replica<R> &s) {  // qa/precompute_nh.cpp:12-20
    for (int idx = 0; idx < s.n; ++idx) {
        int iprev = s.prev[idx], inext = s.next[idx];
        if (iprev < 0 || inext < 0) continue;
        V3<R> v1 = s.r[idx] - s.r[iprev], v2 = s.r[inext] - s.r[idx];
        s.nvec[idx] = unit(v2 - v1);
        s.hvec[idx] = unit(cross(v2, v1));
    }
}
template <typename R> void qa_sift_one(const Params<R> &P, Replica<R> &s, int idx, std::vector<Cand> &out) {
    auto const &node = s.fp[idx];  // qa/sift_candidates.cpp:15-77
    if (node.vacant) return;
    int i1 = node.i1, i2 = node.i2;
    V3<R> r12 = s.box.r_uv(s.r[i1], s.r[i2]);
    R r12_rn = norm_inv(r12);
    V3<R> r12_u = r12 * r12_rn;
    V3<R> n1 = s.nvec[i1], n2 = s.nvec[i2], h1 = s.hvec[i1], h2 = s.hvec[i2];
    bool back1 = (std::abs(dot(h1, r12_u)) > P.min_cos_hr);
    bool back2 = (std::abs(dot(h2, r12_u)) > P.min_cos_hr);
    bool bb_extra = (std::abs(dot(h1, h2)) > P.min_cos_hh);
    bool side1 = (dot(n1, r12_u) < P.max_cos_nr);
    bool side2 = (-dot(n2, r12_u) < P.max_cos_nr);
    int a1 = s.atype[i1], a2 = s.atype[i2];
    int type = 0;
    if (back1 && back2 && bb_extra) type = 1;
    else if (back1 && side2) type = 2;
    else if (side1 && back2) type = 3;
    else if (side1 && side2) type = 4 + a1 + a2 * XMD_NUM_AA;
    if (type == 0) return;
    if (R(1.0f) > r12_rn * P.req_min_dist[type]) return;
    int pt2 = P.ptype[a2], pt1 = P.ptype[a1];
    Sync d1{{(int8_t)back1, (int8_t)side1, (int8_t)(side1 && pt2 == 1), (int8_t)(side1 && pt2 == 2)}};
    if (!sub(s.sync[i1], d1).valid()) return;
    Sync d2{{(int8_t)back2, (int8_t)side2, (int8_t)(side2 && pt1 == 1), (int8_t)(side2 && pt1 == 2)}};
    if (!sub(s.sync[i2], d2).valid()) return;
    out.push_back({i1, i2, idx, (int16_t)type, d1, d2});
}
template <typename R> void qa_process_candidates(Replica<R> &s) {  // qa/process_candidates.cpp:10-29
    for (auto const &c : s.cand) {
        Sync ns1 = sub(s.sync[c.i1], c.d1), ns2 = sub(s.sync[c.i2], c.d2);
        if (ns1.valid() && ns2.valid()) {
            s.cont.push_back({c.i1, c.i2, c.type, 0, s.t, c.d1, c.d2, false});
            s.sync[c.i1] = sub(s.sync[c.i1], c.d1);
            s.sync[c.i2] = sub(s.sync[c.i2], c.d2);
            s.fp[c.fp_idx].vacant = true;
        }
    }
}
// qa/process_contacts.cpp:13-57; `transitions` false evaluates forces only.
template <typename R>
void qa_process_contacts(const Params<R> &P, Replica<R> &s, std::vector<V3<R>> &F, R *V, bool transitions,
                         size_t lo = 0, size_t hi = SIZE_MAX) {
    size_t extent = std::min(hi, s.cont.size());
    for (size_t idx = lo; idx < extent; ++idx) {
        auto &c = s.cont[idx];
        if (c.vacant) continue;
        int i1 = c.i1, i2 = c.i2;
        V3<R> r12 = s.box.r_uv(s.r[i1], s.r[i2]);
        R r12_rn = norm_inv(r12);
        V3<R> r12_u = r12 * r12_rn;
        R saturation = std::min(s.t - c.ref_time, P.cycle_time) * P.cycle_time_inv;
        if (c.status == 1) saturation = R(1.0f) - saturation;
        SinkLJ<R> lj_force = qa_lj(P, c.type);
        R r12_n = r12_rn * norm_sq(r12);
        auto [Vij, dVij_dr] = lj_force(r12_n, r12_rn);
        *V += saturation * Vij;
        V3<R> f = saturation * dVij_dr * r12_u;
        F[i1] += f;
        F[i2] -= f;
        if (!transitions) continue;
        if (c.status == 0 && saturation == R(1.0f)) {
            if (P.factor * lj_force.r_min * r12_rn < R(1.0f)) {
                c.status = 1;
                c.ref_time = s.t;
            }
        } else if (c.status == 1 && saturation == R(0.0f)) {
            c.vacant = true;
            s.sync[i1] = add(s.sync[i1], c.d1);
            s.sync[i2] = add(s.sync[i2], c.d2);
            s.fp.push_back({i1, i2, false});
        }
    }
}

// ------------------------------------------------------------- lists ---
template <typename R> void legacy_update(const Params<R> &P, Replica<R> &s) {  // xmd/src/nl/legacy_update.cpp:6-70 (+D1,D2)
    R pad = P.pad_factor * P.max_cutoff;
    R req_r = P.max_cutoff + pad;
    s.nl.clear();
    size_t nat_idx = 0, ss_idx = 0;
    for (int i1 = 0; i1 < s.n; ++i1) {
        V3<R> r1 = s.r[i1];
        int chain1 = s.chain[i1], seq1 = s.seq[i1];
        for (int i2 = i1 + 1; i2 < s.n; ++i2) {
            int chain2 = s.chain[i2], seq2 = s.seq[i2];
            int diff = seq1 > seq2 ? seq1 - seq2 : seq2 - seq1;
            if (chain1 == chain2 && diff < 3) continue;
            R orig_dist = norm(s.box.r_uv(r1, s.r[i2]));
            if (orig_dist < req_r) {
                std::pair<int, int> p{i1, i2};
                while (nat_idx < s.all_nat_nl.size() && s.all_nat_nl[nat_idx] < p) ++nat_idx;
                bool native = nat_idx < s.all_nat_nl.size() && s.all_nat_nl[nat_idx] == p;
                while (ss_idx < s.all_ss.size() && s.all_ss[ss_idx] < p) ++ss_idx;
                bool ss = ss_idx < s.all_ss.size() && s.all_ss[ss_idx] == p;
                s.nl.push_back({i1, i2, orig_dist, native || ss, ss});
            }
        }
    }
    s.orig_pad = pad;
    s.orig_r = s.r;
    s.orig_box = s.box;
    s.ref_t = s.t;
    s.invalid = false;
}
template <typename R> void update_lists(const Params<R> &P, Replica<R> &s) {
    R pad = s.orig_pad;
    if (P.go_on) {  // xmd/src/forces/go.cpp:8-23
        s.go_cur.clear();
        for (auto const &c : s.all_nat) {
            R cutoff = R(2.0 * c.nat_dist);
            if (norm(s.box.r_uv(s.r[c.i1], s.r[c.i2])) < cutoff + pad) s.go_cur.push_back(c);
        }
    }
    if (P.ss_on) {  // xmd/src/forces/native_ss.cpp:8-19 (+D23: cutoff = harmonic::cutoff(equil))
        s.ss_cur.clear();
        R cutoff = R(2.0) * P.ss_equil;
        for (auto const &c : s.all_ss)
            if (norm(s.box.r_uv(s.r[c.first], s.r[c.second])) < cutoff + pad) s.ss_cur.push_back(c);
    }
    if (P.pauli_on) {  // xmd/src/forces/pauli.cpp:9-24 (+D2,D10)
        s.pauli.clear();
        R min_norm_inv = R(1.0) / (P.pauli_r_excl + pad);
        for (auto const &p : s.nl) {
            if (p.native) continue;
            if (norm_inv(s.box.r_uv(s.r[p.i1], s.r[p.i2])) > min_norm_inv) s.pauli.push_back({p.i1, p.i2});
        }
    }
    if (P.ces_on || P.res_on) {  // xmd/src/forces/es/es_pair.cpp:6-23 (+D2)
        s.es.clear();
        R min_norm_inv = R(1.0) / (P.es_cutoff + pad);
        for (auto const &p : s.nl) {
            if (p.ss) continue;
            if (norm_inv(s.box.r_uv(s.r[p.i1], s.r[p.i2])) > min_norm_inv) {
                R q1_x_q2 = P.q[s.atype[p.i1]] * P.q[s.atype[p.i2]];
                if (q1_x_q2 != 0.0) s.es.push_back({p.i1, p.i2, q1_x_q2});
            }
        }
    }
    if (P.qa_on) {  // xmd/src/forces/qa/free_pair.cpp:6-21 (+D2,D6,D10)
        std::set<std::pair<int, int>> in_contact;
        for (auto const &c : s.cont)
            if (!c.vacant) in_contact.insert({c.i1, c.i2});
        s.fp.clear();
        R min_norm_inv = R(1.0) / (P.max_formation_min_dist + pad);
        for (auto const &p : s.nl) {
            if (p.native) continue;
            if (norm_inv(s.box.r_uv(s.r[p.i1], s.r[p.i2])) > min_norm_inv) {
                if (in_contact.count({p.i1, p.i2})) continue;
                s.fp.push_back({p.i1, p.i2, false});
            }
        }
    }
    if (P.pid_on) {  // xmd/src/forces/pid/bundle.cpp:6-29 (+D2)
        s.pid.clear();
        R min_norm_inv = R(1.0) / (P.pid_cutoff + pad);
        for (auto const &p : s.nl) {
            if (p.native) continue;
            if (norm_inv(s.box.r_uv(s.r[p.i1], s.r[p.i2])) > min_norm_inv) {
                if (s.prev[p.i1] < 0 || s.next[p.i1] < 0 || s.prev[p.i2] < 0 || s.next[p.i2] < 0) continue;
                int16_t type = (int16_t)(s.atype[p.i1] * XMD_NUM_AA + s.atype[p.i2]);
                if (ORACLE_MUTANT == 2) type = (int16_t)(s.atype[p.i2] * XMD_NUM_AA + s.atype[p.i1]);
                s.pid.push_back({p.i1, p.i2, type});
            }
        }
    }
}
// The same list as legacy_update, built by OpenMP threads over i1 (each
// row collected on its own, rows joined in i1 order, membership in the
// native lists by binary search): for the test entry point that rebuilds
// on request.  Stepping keeps legacy_update, serial as in the reference
// (main.cpp:1052-1075 runs it on the master thread).
template <typename R> void legacy_update_par(const Params<R> &P, Replica<R> &s) {
    R pad = P.pad_factor * P.max_cutoff;
    R req_r = P.max_cutoff + pad;
    std::vector<std::vector<NlPair<R>>> rows(s.n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i1 = 0; i1 < s.n; ++i1) {
        V3<R> r1 = s.r[i1];
        int chain1 = s.chain[i1], seq1 = s.seq[i1];
        for (int i2 = i1 + 1; i2 < s.n; ++i2) {
            int chain2 = s.chain[i2], seq2 = s.seq[i2];
            int diff = seq1 > seq2 ? seq1 - seq2 : seq2 - seq1;
            if (chain1 == chain2 && diff < 3) continue;
            R orig_dist = norm(s.box.r_uv(r1, s.r[i2]));
            if (orig_dist < req_r) {
                std::pair<int, int> p{i1, i2};
                bool native = std::binary_search(s.all_nat_nl.begin(), s.all_nat_nl.end(), p);
                bool ss = std::binary_search(s.all_ss.begin(), s.all_ss.end(), p);
                rows[i1].push_back({i1, i2, orig_dist, native || ss, ss});
            }
        }
    }
    s.nl.clear();
    for (auto &row : rows) s.nl.insert(s.nl.end(), row.begin(), row.end());
    s.orig_pad = pad;
    s.orig_r = s.r;
    s.orig_box = s.box;
    s.ref_t = s.t;
    s.invalid = false;
}
template <typename R> void rebuild(const Params<R> &P, Replica<R> &s, bool par = false) {
    if (par) legacy_update_par(P, s);
    else legacy_update(P, s);
    update_lists(P, s);
    s.rebuilds++;
}
template <typename R> void verify(Replica<R> &s) {  // xmd/src/nl/verify.cpp:4-23
    if (s.first_time) {
        s.first_time = false;
        s.invalid = true;
        return;
    }
    R max_r_disp = 0.0f;
    for (int idx = 0; idx < s.n; ++idx) {
        V3<R> dr = s.box.r_uv(s.r[idx], s.orig_r[idx]);
        max_r_disp = std::max(max_r_disp, norm(dr));
    }
    R box_disp = l1_norm(s.box.cell - s.orig_box.cell);
    R total_disp = R(2.0f) * max_r_disp + box_disp;
    if (total_disp > s.orig_pad) s.invalid = true;
}

// -------------------------------------------------------------- langevin ---
template <typename R> void lang_pc(const Params<R> &P, Replica<R> &s) {  // xmd/src/dynamics/lang_pc.cpp:5-41
    RandGen local_gen = s.gen;
    R noise_factor = (R)std::sqrt(2.0 * 1.0 * (double)s.temperature);
    double dt = P.dt;
    double dt_inv = 1.0 / dt;
    R gamma_factor_sqrt = std::sqrt(P.gamma_factor);
    const R mass = R(1.0), mass_inv = R(1.0), mass_rsqrt = R(1.0);
    for (int idx = 0; idx < s.n; ++idx) {
        R gamma = P.gamma_factor * mass;
        R noise_sd = noise_factor * gamma_factor_sqrt * mass_rsqrt;
        auto nxy = local_gen.normal_x2f();
        float noise_z = local_gen.normalf();
        V3<R> noise{(R)nxy.first, (R)nxy.second, (R)noise_z};
        V3<R> a_ = s.F[idx] * mass_inv - gamma * s.v[idx] + noise_sd * noise;
        auto &y0 = s.y[0][idx], &y1 = s.y[1][idx], &y2 = s.y[2][idx], &y3 = s.y[3][idx], &y4 = s.y[4][idx],
             &y5 = s.y[5][idx];
        double h = dt * dt / 2.0;
        V3<double> error{y2.x - (double)a_.x * h, y2.y - (double)a_.y * h, y2.z - (double)a_.z * h};
        y0 -= (3.0 / 16.0) * error;
        y1 -= (251.0 / 360.0) * error;
        y2 -= 1.0 * error;
        y3 -= (11.0 / 18.0) * error;
        y4 -= (1.0 / 6.0) * error;
        y5 -= (1.0 / 60.0) * error;
        y0 += y1 + y2 + y3 + y4 + y5;
        y1 += 2.0 * y2 + 3.0 * y3 + 4.0 * y4 + 5.0 * y5;
        y2 += 3.0 * y3 + 6.0 * y4 + 10.0 * y5;
        y3 += 4.0 * y4 + 10.0 * y5;
        y4 += 5.0 * y5;
        s.r[idx] = {(R)y0.x, (R)y0.y, (R)y0.z};
        V3<double> vv = y1 * dt_inv;
        s.v[idx] = {(R)vv.x, (R)vv.y, (R)vv.z};
    }
    s.true_t += dt;
    s.t = (R)s.true_t;
    s.gen = local_gen;
}

// ----------------------------------------------------------- the system ---
struct SysBase {
    virtual ~SysBase() {}
    int n_rep = 0, n_total = 0;
    std::vector<int> rep_off;
};

template <typename R> struct System : SysBase {
    Params<R> P;
    std::vector<Replica<R>> reps;

    void force_phase(Replica<R> &s, uint32_t mask, std::vector<V3<R>> &F, R *Vt, bool qa_mutate) {
        // order of xmd/src/freeform/main.cpp:1029-1047
        auto on = [&](int t) { return (mask >> t) & 1u; };
        R dummy;
        auto V = [&](int t) -> R * { return Vt ? &Vt[t] : &dummy; };
        if (P.chir_on && on(XMD_TERM_CHIRAL)) eval_chiral(P, s, F, V(XMD_TERM_CHIRAL));
        if (P.tether_on && on(XMD_TERM_TETHER)) eval_tether(P, s, F, V(XMD_TERM_TETHER));
        if (P.ang_on && on(XMD_TERM_NAT_ANG)) eval_angle(P, s, F, V(XMD_TERM_NAT_ANG));
        if (P.cnd_on && on(XMD_TERM_CND)) eval_cnd(P, s, F, V(XMD_TERM_CND));
        if (P.snd_on && on(XMD_TERM_SND)) eval_snd(P, s, F, V(XMD_TERM_SND));
        if (P.hang_on && on(XMD_TERM_HEUR_ANG)) eval_heur_angle(P, s, F, V(XMD_TERM_HEUR_ANG));
        if (P.hdih_on && on(XMD_TERM_HEUR_DIH)) eval_heur_dih(P, s, F, V(XMD_TERM_HEUR_DIH));
        if (P.pauli_on && on(XMD_TERM_PAULI)) eval_pauli(P, s, F, V(XMD_TERM_PAULI));
        if (P.go_on && on(XMD_TERM_GO)) eval_go(P, s, F, V(XMD_TERM_GO));
        if (P.ss_on && on(XMD_TERM_NAT_SS)) eval_nat_ss(P, s, F, V(XMD_TERM_NAT_SS));
        if (P.ces_on && on(XMD_TERM_CONST_ES)) eval_es(P, s, F, V(XMD_TERM_CONST_ES), false);
        if (P.res_on && on(XMD_TERM_REL_ES)) eval_es(P, s, F, V(XMD_TERM_REL_ES), true);
        if (P.qa_on && on(XMD_TERM_QA)) {
            if (qa_mutate) qa_bookkeeping_pre(s);
            qa_process_contacts(P, s, F, V(XMD_TERM_QA), qa_mutate);
        }
        if (P.pid_on && on(XMD_TERM_PID)) eval_pid(P, s, F, V(XMD_TERM_PID));
        // main.cpp:987-1005
        if (P.solid_on && on(XMD_TERM_SOLID_WALL)) eval_solid_walls(P, s, F, V(XMD_TERM_SOLID_WALL));
        if (P.ljw_on && on(XMD_TERM_LJ_WALL)) eval_lj_walls(P, s, F, V(XMD_TERM_LJ_WALL), qa_mutate);
        if (P.vafm_on && on(XMD_TERM_VEL_AFM)) eval_vel_afm(P, s, F);
        if (P.fafm_on && on(XMD_TERM_FORCE_AFM)) eval_force_afm(s, F);
    }
    void qa_bookkeeping_pre(Replica<R> &s) {
        qa_precompute_nh(s);
        s.cand.clear();
        for (int idx = 0; idx < (int)s.fp.size(); ++idx) qa_sift_one(P, s, idx, s.cand);
        qa_process_candidates(s);
    }
    void ensure_lists(Replica<R> &s) {
        if (s.first_time) verify(s);
        if (s.invalid) rebuild(P, s);
    }
    void step_serial(Replica<R> &s) {
        // initial pass of main.cpp:1009-1024
        if (s.first_time) {
            verify(s);
            s.cand.clear();
            if (s.invalid) rebuild(P, s);
        }
        std::fill(s.F.begin(), s.F.end(), V3<R>{0, 0, 0});
        R Vt[XMD_NUM_TERMS] = {};
        force_phase(s, 0xffffffffu, s.F, Vt, true);
        R V = 0;
        for (int k = 0; k < XMD_NUM_TERMS; ++k) V += Vt[k];
        s.V_last = V;
        if (P.lang_on) lang_pc(P, s);
        s.cand.clear();
        verify(s);
        if (s.invalid) rebuild(P, s);
        s.steps++;
    }

    // The reference's OpenMP structure (main.cpp:909-1077), for timing:
    // one team, thread-private F/V, omp-for over every list, atomic
    // reduction, master-only integrator and rebuild.
    void step_omp(Replica<R> &s, int64_t nsteps, int nthreads) {
#ifdef _OPENMP
        if (s.first_time) {
            verify(s);
            if (s.invalid) rebuild(P, s);
        }
        std::vector<R> Vshared(1, 0);
#pragma omp parallel num_threads(nthreads)
        {
            std::vector<V3<R>> tF(s.n);
            for (int64_t st = 0; st < nsteps; ++st) {
                std::fill(tF.begin(), tF.end(), V3<R>{0, 0, 0});
                R tV = 0;
                omp_terms(s, tF, &tV);
                // reduce_vf.cpp:4-11: every thread adds its whole buffer
                // with atomic updates (vec3_ref::atomic_add, vec3.h:93-102)
#pragma omp atomic update
                Vshared[0] += tV;
                for (int i = 0; i < s.n; ++i) {
#pragma omp atomic update
                    s.F[i].x += tF[i].x;
#pragma omp atomic update
                    s.F[i].y += tF[i].y;
#pragma omp atomic update
                    s.F[i].z += tF[i].z;
                }
#pragma omp barrier
#pragma omp master
                {
                    s.V_last = Vshared[0];
                    Vshared[0] = 0;
                    if (P.lang_on) lang_pc(P, s);
                    std::fill(s.F.begin(), s.F.end(), V3<R>{0, 0, 0});
                    s.cand.clear();
                    verify(s);
                    if (s.invalid) rebuild(P, s);
                    s.steps++;
                }
#pragma omp barrier
            }
        }
#else
        for (int64_t st = 0; st < nsteps; ++st) step_serial(s);
        (void)nthreads;
#endif
    }
#ifdef _OPENMP
    // Each thread evaluates its static chunk of every list (the reference's
    // `omp for schedule(static) nowait`) into its private F/V.
    void omp_terms(Replica<R> &s, std::vector<V3<R>> &F, R *V) {
        auto chunk = [&](size_t n, size_t &lo, size_t &hi) {
            int nt = omp_get_num_threads(), id = omp_get_thread_num();
            size_t per = (n + nt - 1) / nt;
            lo = std::min(n, per * id);
            hi = std::min(n, lo + per);
        };
        size_t lo, hi;
        if (P.chir_on) { chunk(s.quads.size(), lo, hi); range_eval(s, &Replica<R>::quads, lo, hi, F, V, 8); }
        if (P.tether_on) { chunk(s.tethers.size(), lo, hi); range_eval(s, &Replica<R>::tethers, lo, hi, F, V, 0); }
        if (P.ang_on) { chunk(s.angles.size(), lo, hi); range_eval(s, &Replica<R>::angles, lo, hi, F, V, 1); }
        if (P.cnd_on || P.snd_on) { chunk(s.dihs.size(), lo, hi); range_eval(s, &Replica<R>::dihs, lo, hi, F, V, 2); }
        if (P.hang_on) { chunk(s.hangles.size(), lo, hi); if (lo < hi) eval_heur_angle(P, s, F, V, lo, hi); }
        if (P.hdih_on) { chunk(s.hdihs.size(), lo, hi); if (lo < hi) eval_heur_dih(P, s, F, V, lo, hi); }
        if (P.pauli_on) { chunk(s.pauli.size(), lo, hi); range_eval(s, &Replica<R>::pauli, lo, hi, F, V, 3); }
        if (P.go_on) { chunk(s.go_cur.size(), lo, hi); range_eval(s, &Replica<R>::go_cur, lo, hi, F, V, 4); }
        if (P.ss_on) { chunk(s.ss_cur.size(), lo, hi); range_eval(s, &Replica<R>::ss_cur, lo, hi, F, V, 9); }
        if (P.ces_on || P.res_on) { chunk(s.es.size(), lo, hi); range_eval(s, &Replica<R>::es, lo, hi, F, V, 5); }
        if (P.qa_on) {
#pragma omp barrier
#pragma omp master
            { qa_bookkeeping_pre(s); }
#pragma omp barrier
            chunk(s.cont.size(), lo, hi);
            range_eval(s, &Replica<R>::cont, lo, hi, F, V, 7);
#pragma omp barrier
#pragma omp master
            { qa_apply_transitions(s); }
#pragma omp barrier
        }
        if (P.pid_on) { chunk(s.pid.size(), lo, hi); range_eval(s, &Replica<R>::pid, lo, hi, F, V, 6); }
        if (P.solid_on) { chunk((size_t)s.n, lo, hi); if (lo < hi) eval_solid_walls(P, s, F, V, lo, hi); }
        if (P.ljw_on) { chunk(s.ljw.size(), lo, hi); if (lo < hi) eval_lj_walls(P, s, F, V, true, lo, hi); }
        if (P.vafm_on) { chunk(s.vtips.size(), lo, hi); if (lo < hi) eval_vel_afm(P, s, F, lo, hi); }
        if (P.fafm_on) { chunk(s.ftips.size(), lo, hi); if (lo < hi) eval_force_afm(s, F, lo, hi); }
    }
    // Contact state transitions of process_contacts, serially, after the
    // parallel force pass (forces of the step are unaffected).
    void qa_apply_transitions(Replica<R> &s) {
        std::vector<V3<R>> Fd(s.n);
        R Vd = 0;
        qa_process_contacts(P, s, Fd, &Vd, true);
    }
    template <typename Item>
    void range_eval(Replica<R> &s, std::vector<Item> Replica<R>::*, size_t lo, size_t hi, std::vector<V3<R>> &F, R *V,
                    int kind) {
        if (lo >= hi) return;
        switch (kind) {
        case 0: eval_tether(P, s, F, V, lo, hi); break;
        case 1: eval_angle(P, s, F, V, lo, hi); break;
        case 2:
            if (P.cnd_on) eval_cnd(P, s, F, V, lo, hi);
            if (P.snd_on) eval_snd(P, s, F, V, lo, hi);
            break;
        case 3: eval_pauli(P, s, F, V, lo, hi); break;
        case 4: eval_go(P, s, F, V, lo, hi); break;
        case 5:
            if (P.ces_on) eval_es(P, s, F, V, false, lo, hi);
            if (P.res_on) eval_es(P, s, F, V, true, lo, hi);
            break;
        case 6: eval_pid(P, s, F, V, lo, hi); break;
        case 7: qa_process_contacts(P, s, F, V, false, lo, hi); break;
        case 8: eval_chiral(P, s, F, V, lo, hi); break;
        case 9: eval_nat_ss(P, s, F, V, lo, hi); break;
        }
    }
#endif
};

// Copy the params struct into `real`, following main.cpp's expressions.
template <typename R> int setup_params(const xmd_params *p, Params<R> &P) {
    P.lang_on = p->langevin_enabled != 0;
    P.dt = p->dt;
    P.gamma_factor = (R)p->gamma_factor;
    P.pad_factor = (R)p->pad_factor;
    P.tether_on = p->tether_enabled;
    P.chir_on = p->chir_enabled;
    P.ang_on = p->nat_ang_enabled;
    P.cnd_on = p->cnd_enabled;
    P.snd_on = p->snd_enabled;
    P.pauli_on = p->pauli_enabled;
    P.go_on = p->go_enabled;
    P.ces_on = p->const_es_enabled;
    P.res_on = p->rel_es_enabled;
    P.qa_on = p->qa_enabled;
    P.pid_on = p->pid_enabled;
    P.ss_on = p->nat_ss_enabled;
    P.solid_on = p->solid_walls_enabled;
    P.ljw_on = p->lj_walls_enabled;
    P.vafm_on = p->vel_afm_enabled;
    P.fafm_on = p->force_afm_enabled;
    P.hang_on = p->heur_ang_enabled;
    P.hdih_on = p->heur_dih_enabled;
    for (int k = 0; k < 9; ++k) {
        for (int d = 0; d < 7; ++d) P.hang_c[k][d] = (R)p->heur_ang_coeffs[k][d];
        for (int d = 0; d < 6; ++d) P.hdih_c[k][d] = (R)p->heur_dih_coeffs[k][d];
    }
    P.solid_eps = (R)p->solid_walls_depth;
    P.solid_cut = (R)p->solid_walls_cutoff;
    P.ljw_min_dist = (R)p->lj_walls_min_dist;
    P.ljw_factor = (R)(p->lj_walls_breaking_factor * std::pow(2.0, -1.0 / 6.0));  // main.cpp:840
    P.ljw_cycle = (R)p->lj_walls_phase_duration;
    P.ljw_cycle_inv = (R)(1.0 / (double)P.ljw_cycle);
    P.vafm_H1 = (R)p->vel_afm_H1;
    P.vafm_H2 = (R)p->vel_afm_H2;
    P.ss_H1 = (R)p->nat_ss_H1;
    P.ss_equil = (R)p->nat_ss_equil_dist;
    P.H1 = (R)p->tether_H1;
    P.H2 = (R)p->tether_H2;
    P.def_len = (R)p->tether_def_length;
    P.e_chi = (R)p->chir_e;
    P.ang_k = (R)p->nat_ang_k;
    P.CDA = (R)p->cnd_CDA;
    P.CDB = (R)p->cnd_CDB;
    P.CDH = (R)p->snd_CDH;
    P.pauli_r_excl = (R)p->pauli_r_excl;
    P.pauli_depth = (R)p->pauli_depth;
    P.go_depth = (R)p->go_depth;
    // main.cpp:619-645
    R perm = (R)p->const_es_permittivity, fac = (R)p->rel_es_factor;
    P.ces_Vf = (R)(1.0 / (4.0 * M_PI * (double)perm));
    P.ces_sdi = (R)(1.0 / p->const_es_screen);
    P.res_Vf = (R)(1.0 / (4.0 * M_PI * (double)fac));
    P.res_sdi = (R)(1.0 / p->rel_es_screen);
    P.es_cutoff = p->const_es_enabled ? (R)(2.0 * p->const_es_screen) : (R)(2.0 * p->rel_es_screen);
    for (int a = 0; a < XMD_NUM_AA; ++a) P.q[a] = (R)p->aa_charge[a];
    // lj_variants.cpp:8-63
    P.bb = {(R)p->lj_bb_depth, (R)p->lj_bb_rmin};
    P.bs = {(R)p->lj_bs_depth, (R)p->lj_bs_rmin};
    P.sb = P.bs;
    for (int k = 0; k < XMD_NUM_SS; ++k) {
        P.ss[k].depth = (R)p->ss_depth[k];
        P.ss[k].r_max = (R)p->ss_rmax[k];
        P.ss[k].r_min = p->ss_sinking ? P.bb.r_min : P.ss[k].r_max;
    }
    // main.cpp:765-812
    P.bbp_lam = {(R)p->pid_bbp_psi0, (R)p->pid_bbp_alpha, p->pid_lambda_version};
    P.bbm_lam = {(R)p->pid_bbm_psi0, (R)p->pid_bbm_alpha, p->pid_lambda_version};
    P.ss_lam = {(R)p->pid_ss_psi0, (R)p->pid_ss_alpha, p->pid_lambda_version};
    P.bbp_lj = {(R)p->pid_bbp_depth, (R)p->pid_bbp_rmin};
    P.bbm_lj = {(R)p->pid_bbm_depth, (R)p->pid_bbm_rmin};
    if (ORACLE_MUTANT == 1) std::swap(P.bbp_lj, P.bbm_lj);
    R pc = 0.0;
    pc = std::max(pc, (R)(2.0 * P.bb.r_min));
    pc = std::max(pc, (R)(2.0 * P.bs.r_min));
    pc = std::max(pc, (R)(2.0 * P.sb.r_min));
    for (int k = 0; k < XMD_NUM_SS; ++k) pc = std::max(pc, (R)(2.0 * P.ss[k].r_max));
    P.pid_cutoff = pc;
    // main.cpp:676-747
    P.cycle_time = (R)p->qa_phase_duration;
    P.cycle_time_inv = (R)(1.0 / P.cycle_time);
    R breaking_factor = (R)p->qa_breaking_factor;
    P.factor = breaking_factor * (R)std::pow(2.0f, -1.0f / 6.0f);
    P.min_cos_hr = (R)p->qa_min_cos_hr;
    P.min_cos_hh = (R)p->qa_min_cos_hh;
    P.max_cos_nr = (R)p->qa_max_cos_nr;
    for (int k = 0; k < XMD_QA_NUM_TYPES; ++k) P.req_min_dist[k] = 0;
    P.req_min_dist[1] = P.bb.r_min;
    P.req_min_dist[2] = P.bs.r_min;
    P.req_min_dist[3] = P.sb.r_min;
    for (int a1 = 0; a1 < XMD_NUM_AA; ++a1)
        for (int a2 = 0; a2 < XMD_NUM_AA; ++a2)
            P.req_min_dist[4 + a1 + a2 * XMD_NUM_AA] = P.ss[a1 * XMD_NUM_AA + a2].r_max;
    R mfmd = 0.0;
    for (int k = 0; k < XMD_QA_NUM_TYPES; ++k) mfmd = std::max(mfmd, P.req_min_dist[k]);
    P.max_formation_min_dist = mfmd;
    for (int a = 0; a < XMD_NUM_AA; ++a) {
        P.ptype[a] = p->aa_polarization[a];
        for (int k = 0; k < 4; ++k) P.limits[a].v[k] = p->aa_limits[a][k];
    }
    return 0;
}

template <typename R> int build_system(const xmd_params *p, const xmd_topology *t, System<R> &S) {
    if (!p || !t) return fail("null params/topology");
    if (t->n_replicas <= 0 || t->n_beads <= 0) return fail("empty topology");
    if (t->replica_offset[0] != 0 || t->replica_offset[t->n_replicas] != t->n_beads)
        return fail("replica_offset must span [0, n_beads]");
    setup_params(p, S.P);
    auto &P = S.P;
    // max_cutoff aggregation, main.cpp:500,551,630,644,746,811
    R mc = 0.0;
    if (P.pauli_on) mc = std::max(mc, P.pauli_r_excl);
    if (P.go_on)
        for (int k = 0; k < t->n_native; ++k) mc = std::max(mc, (R)(2.0 * (double)(R)t->nat_dist[k]));
    if (P.ces_on) mc = std::max(mc, (R)(2.0 * p->const_es_screen));  // main.cpp:628
    if (P.res_on) mc = std::max(mc, (R)(2.0 * p->rel_es_screen));    // main.cpp:642
    if (P.qa_on) mc = std::max(mc, P.max_formation_min_dist);
    if (P.pid_on) mc = std::max(mc, P.pid_cutoff);
    if (P.ss_on) mc = std::max(mc, R(2.0) * P.ss_equil);  // main.cpp:592-593
    if (p->max_cutoff_override > 0) mc = (R)p->max_cutoff_override;
    P.max_cutoff = mc;

    for (int w = 0; w < t->n_solid_walls; ++w) {
        std::array<R, 6> a;
        for (int k = 0; k < 6; ++k) a[k] = (R)t->solid_walls[6 * w + k];
        P.solid_walls.push_back(a);
    }
    for (int w = 0; w < t->n_lj_walls; ++w) {
        std::array<R, 6> a;
        for (int k = 0; k < 6; ++k) a[k] = (R)t->lj_walls[6 * w + k];
        P.lj_walls.push_back(a);
    }
    S.n_rep = t->n_replicas;
    S.n_total = t->n_beads;
    S.rep_off.assign(t->replica_offset, t->replica_offset + t->n_replicas + 1);
    S.reps.resize(S.n_rep);
    std::vector<int> chain_of(t->n_beads, -1), seq_of(t->n_beads, 0);
    for (int c = 0; c < t->n_chains; ++c)
        for (int i = t->chain_offset[c]; i < t->chain_offset[c + 1]; ++i) {
            chain_of[i] = c;
            seq_of[i] = i - t->chain_offset[c];
        }
    for (int i = 0; i < t->n_beads; ++i)
        if (chain_of[i] < 0) return fail("bead outside every chain");
    for (int k = 0; k < S.n_rep; ++k) {
        auto &s = S.reps[k];
        int off = t->replica_offset[k], n = t->replica_offset[k + 1] - off;
        if (n <= 0) return fail("empty replica");
        s.n = n;
        s.off = off;
        s.atype.resize(n);
        s.prev.assign(n, -1);
        s.next.assign(n, -1);
        s.chain.resize(n);
        s.seq.resize(n);
        s.r.resize(n);
        s.v.assign(n, {0, 0, 0});
        s.F.assign(n, {0, 0, 0});
        for (auto &yy : s.y) yy.assign(n, {0, 0, 0});
        for (int i = 0; i < n; ++i) {
            int g = off + i;
            s.atype[i] = t->atype[g];
            if (s.atype[i] < 0 || s.atype[i] >= XMD_NUM_AA) return fail("atype out of range");
            s.chain[i] = chain_of[g];
            s.seq[i] = seq_of[g];
            if (g > 0 && chain_of[g - 1] == chain_of[g]) s.prev[i] = i - 1;
            if (g + 1 < t->n_beads && chain_of[g + 1] == chain_of[g]) s.next[i] = i + 1;
            if ((s.prev[i] >= 0 && g - 1 < off) || (s.next[i] >= 0 && g + 1 >= off + n))
                return fail("chain spans replicas");
            s.r[i] = {(R)t->r0[g], (R)t->r0[t->n_beads + g], (R)t->r0[2 * t->n_beads + g]};
            s.y[0][i] = {t->r0[g], t->r0[t->n_beads + g], t->r0[2 * t->n_beads + g]};
        }
        s.box.set_cell({(R)t->box[3 * k], (R)t->box[3 * k + 1], (R)t->box[3 * k + 2]});
        s.temperature = (R)t->temperature[k];
        // walls (the same planes in every replica) and AFM tips of this replica
        s.ljw.assign((size_t)n * P.lj_walls.size(), {});
        for (int c = 0; c < t->n_vel_afm; ++c) {
            int g = t->vel_afm_idx[c];
            if (g < off || g >= off + n) continue;
            const double *o = t->vel_afm_orig + 3 * c, *v = t->vel_afm_v + 3 * c;
            V3<R> orig = std::isnan(o[0]) ? s.r[g - off] : V3<R>{(R)o[0], (R)o[1], (R)o[2]};
            s.vtips.push_back({g - off, orig, {(R)v[0], (R)v[1], (R)v[2]}});
        }
        for (int c = 0; c < t->n_force_afm; ++c) {
            int g = t->force_afm_idx[c];
            if (g < off || g >= off + n) continue;
            const double *f = t->force_afm_F + 3 * c;
            s.ftips.push_back({g - off, {(R)f[0], (R)f[1], (R)f[2]}});
        }
        s.gen.state = t->rng_state[k];
        s.nvec.assign(n, {0, 0, 0});
        s.hvec.assign(n, {0, 0, 0});
        s.sync.resize(n);
        for (int i = 0; i < n; ++i) s.sync[i] = P.limits[s.atype[i]];
        // native contacts of this replica (sorted by the caller)
        for (int c = 0; c < t->n_native; ++c) {
            int a = t->nat_i1[c], b = t->nat_i2[c];
            if (a < off || a >= off + n) continue;
            if (b < off || b >= off + n || !(a < b)) return fail("native contact crosses replicas or i1 >= i2");
            s.all_nat.push_back({a - off, b - off, (R)t->nat_dist[c]});
            s.all_nat_nl.push_back({a - off, b - off});
        }
        if (!std::is_sorted(s.all_nat_nl.begin(), s.all_nat_nl.end())) return fail("native contacts not sorted");
        // disulfide bonds of this replica (main.cpp:560-577)
        for (int c = 0; c < t->n_ssbonds; ++c) {
            int a = t->ss_i1[c], b = t->ss_i2[c];
            if (a < off || a >= off + n) continue;
            if (b < off || b >= off + n || !(a < b)) return fail("ssbond crosses replicas or i1 >= i2");
            s.all_ss.push_back({a - off, b - off});
        }
        if (!std::is_sorted(s.all_ss.begin(), s.all_ss.end())) return fail("ssbonds not sorted");
        for (auto const &q : s.all_ss)
            if (std::binary_search(s.all_nat_nl.begin(), s.all_nat_nl.end(), q))
                return fail("pair is both a native contact and an ssbond");
        // bonded lists: consecutive beads of each chain (model tethers,
        // angles, dihedrals), main.cpp:377-479
        for (int i = 0; i < n; ++i) {
            int g = off + i;
            if (s.next[i] >= 0) {
                double L = t->tether_len ? t->tether_len[g] : NAN;
                R nat = std::isnan(L) ? P.def_len : (R)L;
                s.tethers.push_back({i, i + 1, nat});
            }
            bool has3 = s.next[i] >= 0 && s.next[i + 1] >= 0;
            bool has4 = has3 && s.next[i + 2] >= 0;
            if (has3 && t->nat_theta && !std::isnan(t->nat_theta[g])) s.angles.push_back({i, i + 1, i + 2, (R)t->nat_theta[g]});
            if (has4 && t->nat_phi && !std::isnan(t->nat_phi[g])) s.dihs.push_back({i, i + 1, i + 2, i + 3, (R)t->nat_phi[g]});
            if (has4 && t->nat_chir && !std::isnan(t->nat_chir[g]))
                s.quads.push_back({i, i + 1, i + 2, i + 3, (R)t->nat_chir[g], (R)t->nat_chir_factor[g]});
            // heuristic entries (D28): model angles / dihedrals without a
            // native value, typed by beads 1, 2 (angle) and 2, 3 (dihedral)
            uint8_t hb = t->heur_bonded ? t->heur_bonded[g] : 0;
            if ((hb & 1) && !has3) return fail("heuristic angle does not join consecutive residues");
            if ((hb & 2) && !has4) return fail("heuristic dihedral does not join consecutive residues");
            auto cls = [&](int b) { return s.atype[b] == 7 ? 0 : (s.atype[b] == 14 ? 1 : 2); };  // GLY, PRO
            if ((hb & 1) && !(t->nat_theta && !std::isnan(t->nat_theta[g])))
                s.hangles.push_back({i, i + 1, i + 2, 3 * cls(i) + cls(i + 1)});
            if ((hb & 2) && !(t->nat_phi && !std::isnan(t->nat_phi[g])))
                s.hdihs.push_back({i, i + 1, i + 2, i + 3, 3 * cls(i + 1) + cls(i + 2)});
        }
    }
    return 0;
}

SysBase *as_base(oracle_sys *s) { return reinterpret_cast<SysBase *>(s); }

template <typename Fn> int dispatch(oracle_sys *h, Fn fn) {
    SysBase *b = as_base(h);
    if (!b) return fail("null handle");
    if (auto *d = dynamic_cast<System<double> *>(b)) return fn(*d);
    if (auto *f = dynamic_cast<System<float> *>(b)) return fn(*f);
    return fail("bad handle");
}

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

int oracle_create(const xmd_params *p, const xmd_topology *t, int real_kind, oracle_sys **out) {
    SysBase *b;
    int rc;
    if (real_kind == 0) {
        auto *s = new System<double>();
        rc = build_system(p, t, *s);
        b = s;
    } else {
        auto *s = new System<float>();
        rc = build_system(p, t, *s);
        b = s;
    }
    if (rc) {
        delete b;
        return rc;
    }
    *out = reinterpret_cast<oracle_sys *>(b);
    return 0;
}

void oracle_destroy(oracle_sys *s) { delete as_base(s); }

int oracle_set_state(oracle_sys *h, int replica, const double *y, uint64_t rng, double t) {
    return dispatch(h, [&](auto &S) {
        using R = std::remove_reference_t<decltype(S.reps[0].t)>;
        if (replica < 0 || replica >= S.n_rep) return fail("replica out of range");
        auto &s = S.reps[replica];
        int n = s.n;
        for (int k = 0; k < 6; ++k)
            for (int i = 0; i < n; ++i) s.y[k][i] = {y[(3 * k) * n + i], y[(3 * k + 1) * n + i], y[(3 * k + 2) * n + i]};
        for (int i = 0; i < n; ++i) {
            s.r[i] = {(R)s.y[0][i].x, (R)s.y[0][i].y, (R)s.y[0][i].z};
            V3<double> vv = s.y[1][i] * (1.0 / S.P.dt);
            s.v[i] = {(R)vv.x, (R)vv.y, (R)vv.z};
        }
        s.gen.state = rng;
        s.true_t = t;
        s.t = (R)t;
        s.first_time = true;
        s.invalid = true;
        return 0;
    });
}

int oracle_get_state(oracle_sys *h, int replica, double *y, double *v, uint64_t *rng, double *t) {
    return dispatch(h, [&](auto &S) {
        if (replica < 0 || replica >= S.n_rep) return fail("replica out of range");
        auto &s = S.reps[replica];
        int n = s.n;
        if (y)
            for (int k = 0; k < 6; ++k)
                for (int i = 0; i < n; ++i) {
                    y[(3 * k) * n + i] = s.y[k][i].x;
                    y[(3 * k + 1) * n + i] = s.y[k][i].y;
                    y[(3 * k + 2) * n + i] = s.y[k][i].z;
                }
        if (v)
            for (int i = 0; i < n; ++i) {
                v[i] = s.v[i].x;
                v[n + i] = s.v[i].y;
                v[2 * n + i] = s.v[i].z;
            }
        if (rng) *rng = s.gen.state;
        if (t) *t = s.true_t;
        return 0;
    });
}

int oracle_set_positions(oracle_sys *h, int replica, const double *r) {
    return dispatch(h, [&](auto &S) {
        using R = std::remove_reference_t<decltype(S.reps[0].t)>;
        if (replica < 0 || replica >= S.n_rep) return fail("replica out of range");
        auto &s = S.reps[replica];
        int n = s.n;
        for (int i = 0; i < n; ++i) {
            s.y[0][i] = {r[i], r[n + i], r[2 * n + i]};
            s.r[i] = {(R)r[i], (R)r[n + i], (R)r[2 * n + i]};
        }
        s.first_time = true;
        s.invalid = true;
        return 0;
    });
}

int oracle_rebuild_lists(oracle_sys *h) {
    return dispatch(h, [&](auto &S) {
        for (auto &s : S.reps) {
            s.first_time = false;
            rebuild(S.P, s, s.n >= 4096);
        }
        return 0;
    });
}

int oracle_get_list(oracle_sys *h, int replica, int id, int32_t *i1, int32_t *i2, int64_t cap, int64_t *n) {
    return dispatch(h, [&](auto &S) {
        if (replica < 0 || replica >= S.n_rep) return fail("replica out of range");
        auto &s = S.reps[replica];
        std::vector<std::pair<int, int>> out;
        switch (id) {
        case XMD_LIST_NL_ALL:
            for (auto &p : s.nl) out.push_back({p.i1, p.i2});
            break;
        case XMD_LIST_NL_NATIVE:
            for (auto &p : s.nl)
                if (p.native && !p.ss) out.push_back({p.i1, p.i2});
            break;
        case XMD_LIST_NL_NAT_SS:
            for (auto &p : s.nl)
                if (p.ss) out.push_back({p.i1, p.i2});
            break;
        case XMD_LIST_NAT_SS: out = s.ss_cur; break;
        case XMD_LIST_NL_NON_NATIVE:
            for (auto &p : s.nl)
                if (!p.native) out.push_back({p.i1, p.i2});
            break;
        case XMD_LIST_GO:
            for (auto &p : s.go_cur) out.push_back({p.i1, p.i2});
            break;
        case XMD_LIST_PAULI: out = s.pauli; break;
        case XMD_LIST_ES:
            for (auto &p : s.es) out.push_back({p.i1, p.i2});
            break;
        case XMD_LIST_PID:
            for (auto &p : s.pid) out.push_back({p.i1, p.i2});
            break;
        case XMD_LIST_QA_FREE:
            for (auto &p : s.fp)
                if (!p.vacant) out.push_back({p.i1, p.i2});
            break;
        default: return fail("bad list id");
        }
        *n = (int64_t)out.size();
        if (i1 && i2)
            for (int64_t k = 0; k < std::min<int64_t>(cap, *n); ++k) {
                i1[k] = out[k].first;
                i2[k] = out[k].second;
            }
        return 0;
    });
}

int oracle_get_wall_state(oracle_sys *h, int replica, int8_t *status, double *joint, double *ref_t) {
    return dispatch(h, [&](auto &S) {
        if (replica < 0 || replica >= S.n_rep) return fail("replica out of range");
        auto &s = S.reps[replica];
        for (size_t k = 0; k < s.ljw.size(); ++k) {
            if (status) status[k] = s.ljw[k].status;
            if (joint) joint[3 * k] = s.ljw[k].joint.x, joint[3 * k + 1] = s.ljw[k].joint.y, joint[3 * k + 2] = s.ljw[k].joint.z;
            if (ref_t) ref_t[k] = s.ljw[k].ref_t;
        }
        return 0;
    });
}

int oracle_set_wall_state(oracle_sys *h, int replica, const int8_t *status, const double *joint, const double *ref_t) {
    return dispatch(h, [&](auto &S) {
        using R = std::remove_reference_t<decltype(S.reps[0].t)>;
        if (replica < 0 || replica >= S.n_rep) return fail("replica out of range");
        auto &s = S.reps[replica];
        for (size_t k = 0; k < s.ljw.size(); ++k) {
            if (status[k] < 0 || status[k] > 2) return fail("bad wall status");
            s.ljw[k].status = status[k];
            s.ljw[k].joint = {(R)joint[3 * k], (R)joint[3 * k + 1], (R)joint[3 * k + 2]};
            s.ljw[k].ref_t = (R)ref_t[k];
        }
        return 0;
    });
}

int oracle_eval_forces(oracle_sys *h, uint32_t mask, double *F, double *V_terms) {
    return dispatch(h, [&](auto &S) {
        using R = std::remove_reference_t<decltype(S.reps[0].t)>;
        for (int k = 0; k < S.n_rep; ++k) {
            auto &s = S.reps[k];
            S.ensure_lists(s);
            std::vector<V3<R>> Fl(s.n, V3<R>{0, 0, 0});
            R Vt[XMD_NUM_TERMS] = {};
            S.force_phase(s, mask, Fl, Vt, false);
            for (int i = 0; i < s.n; ++i) {
                F[s.off + i] = Fl[i].x;
                F[S.n_total + s.off + i] = Fl[i].y;
                F[2 * S.n_total + s.off + i] = Fl[i].z;
            }
            if (V_terms)
                for (int tt = 0; tt < XMD_NUM_TERMS; ++tt) V_terms[k * XMD_NUM_TERMS + tt] = Vt[tt];
        }
        return 0;
    });
}

int oracle_qa_update(oracle_sys *h) {
    return dispatch(h, [&](auto &S) {
        using R = std::remove_reference_t<decltype(S.reps[0].t)>;
        if (!S.P.qa_on) return 0;
        for (auto &s : S.reps) {
            S.ensure_lists(s);
            std::vector<V3<R>> Fl(s.n, V3<R>{0, 0, 0});
            R V = 0;
            S.qa_bookkeeping_pre(s);
            qa_process_contacts(S.P, s, Fl, &V, true);
            s.cand.clear();
        }
        return 0;
    });
}

int oracle_get_qa_state(oracle_sys *h, int replica, xmd_qa_contact *c, int64_t cap, int64_t *n, int8_t *sync) {
    return dispatch(h, [&](auto &S) {
        if (replica < 0 || replica >= S.n_rep) return fail("replica out of range");
        auto &s = S.reps[replica];
        int64_t k = 0;
        for (auto &ct : s.cont) {
            if (ct.vacant) continue;
            if (c && k < cap) {
                xmd_qa_contact &o = c[k];
                o.i1 = ct.i1;
                o.i2 = ct.i2;
                o.type = ct.type;
                o.status = ct.status;
                o._pad = 0;
                std::memcpy(o.diff1, ct.d1.v, 4);
                std::memcpy(o.diff2, ct.d2.v, 4);
                o.ref_time = ct.ref_time;
            }
            ++k;
        }
        *n = k;
        if (sync)
            for (int i = 0; i < s.n; ++i) std::memcpy(sync + 4 * i, s.sync[i].v, 4);
        return 0;
    });
}

int oracle_set_qa_state(oracle_sys *h, int replica, const xmd_qa_contact *c, int64_t n, const int8_t *sync) {
    return dispatch(h, [&](auto &S) {
        using R = std::remove_reference_t<decltype(S.reps[0].t)>;
        if (replica < 0 || replica >= S.n_rep) return fail("replica out of range");
        auto &s = S.reps[replica];
        s.cont.clear();
        for (int64_t k = 0; k < n; ++k) {
            Contact<R> ct;
            ct.i1 = c[k].i1;
            ct.i2 = c[k].i2;
            ct.type = c[k].type;
            ct.status = c[k].status;
            ct.ref_time = (R)c[k].ref_time;
            std::memcpy(ct.d1.v, c[k].diff1, 4);
            std::memcpy(ct.d2.v, c[k].diff2, 4);
            ct.vacant = false;
            s.cont.push_back(ct);
        }
        if (sync)
            for (int i = 0; i < s.n; ++i) std::memcpy(s.sync[i].v, sync + 4 * i, 4);
        s.invalid = true;
        return 0;
    });
}

int oracle_step(oracle_sys *h, int64_t nsteps, int nthreads) {
    return dispatch(h, [&](auto &S) {
        for (auto &s : S.reps) {
            if (nthreads > 0)
                S.step_omp(s, nsteps, nthreads);
            else
                for (int64_t k = 0; k < nsteps; ++k) S.step_serial(s);
        }
        return 0;
    });
}

int oracle_get_energy(oracle_sys *h, double *V) {
    return dispatch(h, [&](auto &S) {
        for (int k = 0; k < S.n_rep; ++k) V[k] = S.reps[k].V_last;
        return 0;
    });
}

int oracle_get_counters(oracle_sys *h, int64_t *out) {
    return dispatch(h, [&](auto &S) {
        int64_t steps = 0, reb = 0, nl = 0, qa = 0;
        for (auto &s : S.reps) {
            steps = std::max(steps, s.steps);
            reb += s.rebuilds;
            nl += 2 * (int64_t)s.nl.size();
            for (auto &c : s.cont) qa += !c.vacant;
        }
        out[0] = steps;
        out[1] = reb;
        out[2] = nl;
        out[3] = nl;
        out[4] = qa;
        return 0;
    });
}

uint64_t oracle_rng_next(uint64_t *state) {
    RandGen g{*state};
    uint64_t v = g.next();
    *state = g.state;
    return v;
}

float oracle_rng_uniformf(uint64_t *state) {
    RandGen g{*state};
    float v = g.uniformf();
    *state = g.state;
    return v;
}

void oracle_rand_normals(uint64_t state, int64_t n, float *out, uint64_t *state_out) {
    RandGen g{state};  // lang_pc.cpp:15-16
    for (int64_t k = 0; k < n; ++k) {
        auto xy = g.normal_x2f();
        float z = g.normalf();
        out[3 * k] = xy.first;
        out[3 * k + 1] = xy.second;
        out[3 * k + 2] = z;
    }
    if (state_out) *state_out = g.state;
}

// xmd/src/model/model.cpp:92-174: chains laid out as random walks; the
// rotations are applied as Rodrigues matrices R(axis, angle).
static void rodrigues(const double ax[3], double ang, double R[3][3]) {
    double c = std::cos(ang), s = std::sin(ang), t = 1.0 - c;
    double x = ax[0], y = ax[1], z = ax[2];
    R[0][0] = c + x * x * t, R[0][1] = x * y * t - z * s, R[0][2] = x * z * t + y * s;
    R[1][0] = y * x * t + z * s, R[1][1] = c + y * y * t, R[1][2] = y * z * t - x * s;
    R[2][0] = z * x * t - y * s, R[2][1] = z * y * t + x * s, R[2][2] = c + z * z * t;
}
void oracle_morph_into_saw(uint64_t *state, int n_chains, const int *chain_len, const double *orig, double bond,
                           double density, int infer_box, double *pos, double *box) {
    RandGen g{*state};
    int n = 0;
    for (int c = 0; c < n_chains; ++c) n += chain_len[c];
    const double a = std::cbrt((double)n / density);
    int base = 0;
    for (int c = 0; c < n_chains; ++c) {
        double p[3], d[3];
        for (int k = 0; k < 3; ++k) p[k] = g.uniformd(-a / 2.0, a / 2.0);
        g.sphere(d);
        for (int i = 0; i < chain_len[c]; ++i) {
            double nx[3] = {p[0], p[1], p[2]};
            if (i + 1 < chain_len[c]) {
                double bl = bond;
                if (!(bond > 0.0)) {
                    const double *q1 = orig + 3 * (base + i), *q2 = orig + 3 * (base + i + 1);
                    bl = std::sqrt((q2[0] - q1[0]) * (q2[0] - q1[0]) + (q2[1] - q1[1]) * (q2[1] - q1[1]) +
                                   (q2[2] - q1[2]) * (q2[2] - q1[2]));
                }
                for (int k = 0; k < 3; ++k) nx[k] += d[k] * bl;
                double spread = g.uniformd(-M_PI / 12.0, M_PI / 12.0);
                // perpendicular: the larger of d x e_x, d x e_y, normalised
                double px[3] = {0.0, d[2], -d[1]}, py[3] = {-d[2], 0.0, d[0]};
                double nxn = std::sqrt(px[1] * px[1] + px[2] * px[2]), nyn = std::sqrt(py[0] * py[0] + py[2] * py[2]);
                double perp[3];
                for (int k = 0; k < 3; ++k) perp[k] = nxn > nyn ? px[k] / nxn : py[k] / nyn;
                double around = g.uniformd(-M_PI, M_PI);
                double Rs[3][3], Ra[3][3], t[3], u[3];
                rodrigues(perp, spread, Rs);
                rodrigues(d, around, Ra);
                for (int r = 0; r < 3; ++r) t[r] = Rs[r][0] * d[0] + Rs[r][1] * d[1] + Rs[r][2] * d[2];
                for (int r = 0; r < 3; ++r) u[r] = Ra[r][0] * t[0] + Ra[r][1] * t[1] + Ra[r][2] * t[2];
                double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                for (int k = 0; k < 3; ++k) d[k] = u[k] / un;
            }
            for (int k = 0; k < 3; ++k) pos[3 * (base + i) + k] = p[k];
            for (int k = 0; k < 3; ++k) p[k] = nx[k];
        }
        base += chain_len[c];
    }
    if (infer_box) {
        // every extent from x; maxima start at the smallest positive double
        double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::min();
        for (int i = 0; i < n; ++i) lo = std::min(lo, pos[3 * i]), hi = std::max(hi, pos[3 * i]);
        box[0] = box[1] = box[2] = hi - lo;
    }
    *state = g.state;
}

void oracle_prim(int kind, const double *a, double *out) {
    std::pair<double, double> r{0, 0};
    switch (kind) {
    case 0: r = LJ<double>{a[0], a[1]}(a[2]); break;                 // depth, r_min, r_inv
    case 1: r = ShiftedLJ<double>{a[0], a[1]}(a[2]); break;          // depth, r_min, r_inv
    case 2: r = SinkLJ<double>{a[0], a[1], a[2]}(a[3], a[4]); break;  // depth, r_min, r_max, r, r_inv
    case 3: r = Harmonic<double>{a[0], a[1], a[2]}(a[3]); break;     // H1, H2, nat_r, r
    case 4: r = Lambda<double>{a[0], a[1], 0}(a[2]); break;          // cosine: psi_0, alpha, psi
    case 5: r = Lambda<double>{a[0], a[1], 1}(a[2]); break;          // algebraic
    }
    out[0] = r.first;
    out[1] = r.second;
}

}  // extern "C"
