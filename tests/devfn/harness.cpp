// Host build of the engine's per-item physics (xmd_amd/csrc/xmd_dev.hpp) for
// tests: evaluates the same functions in float and double on given
// geometries so the mixed-precision error can be measured on the CPU.
// Test infrastructure only; the engine never loads this library.
#include "../../xmd_amd/csrc/xmd_dev.hpp"

using namespace xmd_dev;

template <typename T> static TermParams<T> params_of(const double *q) {
    TermParams<T> P;
    P.H1 = (T)q[0]; P.H2 = (T)q[1]; P.e_chi = (T)q[2]; P.ang_k = (T)q[3];
    P.CDA = (T)q[4]; P.CDB = (T)q[5]; P.CDH = (T)q[6];
    P.pauli_r_excl = (T)q[7]; P.pauli_depth = (T)q[8]; P.go_depth = (T)q[9];
    P.es_Vf = (T)q[10]; P.es_sdi = (T)q[11]; P.es_relative = (int)q[12];
    P.bbp_psi0 = (T)q[13]; P.bbp_alpha = (T)q[14]; P.bbp_rmin = (T)q[15]; P.bbp_depth = (T)q[16];
    P.bbm_psi0 = (T)q[17]; P.bbm_alpha = (T)q[18]; P.bbm_rmin = (T)q[19]; P.bbm_depth = (T)q[20];
    P.ss_psi0 = (T)q[21]; P.ss_alpha = (T)q[22]; P.lambda_version = (int)q[23];
    P.set_derived();
    return P;
}
template <typename T> static v3<T> V(const double *g) { return {(T)g[0], (T)g[1], (T)g[2]}; }

template <typename T> static void pid_all(int n, const double *geo, const double *ss, const double *prm, double *out) {
    TermParams<T> P = params_of<T>(prm);
    for (int i = 0; i < n; ++i) {
        const double *g = geo + 27 * i;
        v3<T> o[6];
        // geometry: side 1 bonds (g, g+3, g+6), side 2 bonds (g+12, g+15, g+18),
        // r12 at g+24 (g+9 and g+21 repeat r12 and -r12)
        T v = pid_bundle<T>(P, (T)ss[3 * i], (T)ss[3 * i + 1], (T)ss[3 * i + 2],
                            pid_side<T>(V<double>(g), V<double>(g + 3), V<double>(g + 6)),
                            pid_side<T>(V<double>(g + 12), V<double>(g + 15), V<double>(g + 18)),
                            V<double>(g + 24), o);
        for (int k = 0; k < 6; ++k) {
            out[19 * i + 3 * k] = o[k].x;
            out[19 * i + 3 * k + 1] = o[k].y;
            out[19 * i + 3 * k + 2] = o[k].z;
        }
        out[19 * i + 18] = v;
    }
}
template <typename T> static void quad_all(int kind, int n, const double *geo, const double *nat, const double *prm,
                                           double *out) {
    TermParams<T> P = params_of<T>(prm);
    for (int i = 0; i < n; ++i) {
        const double *g = geo + 9 * i;
        v3<T> o[4];
        T v;
        if (kind == 2) v = chiral<T>(V<T>(g), V<T>(g + 3), V<T>(g + 6), P.e_chi, (T)nat[2 * i], (T)nat[2 * i + 1], o);
        else v = nat_dihedral<T>(dih_frame<T>(V<double>(g), V<double>(g + 3), V<double>(g + 6)), kind, P.CDA, P.CDB, P.CDH, (T)nat[i], o);
        for (int k = 0; k < 4; ++k) {
            out[13 * i + 3 * k] = o[k].x;
            out[13 * i + 3 * k + 1] = o[k].y;
            out[13 * i + 3 * k + 2] = o[k].z;
        }
        out[13 * i + 12] = v;
    }
}
template <typename T> static void angle_all(int n, const double *geo, const double *nat, const double *prm,
                                            double *out) {
    TermParams<T> P = params_of<T>(prm);
    for (int i = 0; i < n; ++i) {
        const double *g = geo + 6 * i;
        v3<T> o[3];
        T v = nat_angle<T>(V<double>(g), V<double>(g + 3), P.ang_k, (T)nat[i], o);
        for (int k = 0; k < 3; ++k) {
            out[10 * i + 3 * k] = o[k].x;
            out[10 * i + 3 * k + 1] = o[k].y;
            out[10 * i + 3 * k + 2] = o[k].z;
        }
        out[10 * i + 9] = v;
    }
}
template <typename T> static void pair_all(int kind, int n, const double *d, const double *par, const double *prm,
                                           double *out) {
    TermParams<T> P = params_of<T>(prm);
    for (int i = 0; i < n; ++i) {
        FV<T> f;
        v3<T> dv = V<T>(d + 3 * i);
        if (kind == 0) {  // the engine evaluates tethers in FP64 in both forms
            v3<double> o[2];
            double v = tether<double>(V<double>(d + 3 * i), (double)P.H1, (double)P.H2, par[i], o);
            f.F = {(T)o[0].x, (T)o[0].y, (T)o[0].z};
            f.V = (T)v;
        } else if (kind == 1) f = go_pair<T>(dv, P.go_depth, P.go_d12, (T)par[i]);
        else if (kind == 2) f = pauli_pair<T>(dv, P.pauli_depth, P.pauli_d12, P.pauli_r_excl);
        else f = es_pair<T>(dv, (T)par[i], P.es_Vf, P.es_sdi, P.es_relative);
        out[4 * i] = f.F.x;
        out[4 * i + 1] = f.F.y;
        out[4 * i + 2] = f.F.z;
        out[4 * i + 3] = f.V;
    }
}

extern "C" {
void h_pid(int prec, int n, const double *geo, const double *ss, const double *prm, double *out) {
    if (prec) pid_all<float>(n, geo, ss, prm, out);
    else pid_all<double>(n, geo, ss, prm, out);
}
void h_quad(int prec, int kind, int n, const double *geo, const double *nat, const double *prm, double *out) {
    if (prec) quad_all<float>(kind, n, geo, nat, prm, out);
    else quad_all<double>(kind, n, geo, nat, prm, out);
}
void h_angle(int prec, int n, const double *geo, const double *nat, const double *prm, double *out) {
    if (prec) angle_all<float>(n, geo, nat, prm, out);
    else angle_all<double>(n, geo, nat, prm, out);
}
void h_pair(int prec, int kind, int n, const double *d, const double *par, const double *prm, double *out) {
    if (prec) pair_all<float>(kind, n, d, par, prm, out);
    else pair_all<double>(kind, n, d, par, prm, out);
}
void h_atan2f(int n, const float *y, const float *x, float *out) {
    for (int i = 0; i < n; ++i) out[i] = atan2_f(y[i], x[i]);
}
// lambda value and derivative at psi[i]; out[2i], out[2i+1]
void h_lam(int prec, int version, double psi0, double alpha, int n, const double *psi, double *out) {
    for (int i = 0; i < n; ++i) {
        if (prec) {
            VdV<float> l = lam_eval<float>((float)psi0, (float)alpha, version, (float)psi[i]);
            out[2 * i] = l.V, out[2 * i + 1] = l.dV;
        } else {
            VdV<double> l = lam_eval<double>(psi0, alpha, version, psi[i]);
            out[2 * i] = l.V, out[2 * i + 1] = l.dV;
        }
    }
}
}
