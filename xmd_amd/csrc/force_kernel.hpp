// force_kernel.hpp -- the per-bead force kernel (included by forces.hip
// only, which is compiled with -ffp-contract=fast-honor-pragmas: the force
// arithmetic may fuse multiply-adds, except in blocks that turn contraction
// off; list, QA and noise code elsewhere keeps exact evaluation order).
#pragma once
#include "common.hpp"

namespace xmd_k {

// Force on every slot from every enabled term.  LPB consecutive lanes share
// a bead: its bonded contributions (computed once per quad, angle and bond
// by the block, bonded_phase), native contacts and neighbour-row entries are
// strided over the lanes, then a shuffle tree (fixed order, hence
// deterministic) sums the lanes' partial forces.  The result goes to Fbuf
// (plus the PID spill onto chain neighbours); the integrator runs in
// k_integrate.
template <typename T>
__device__ inline void add3(double &x, double &y, double &z, v3<T> f) {
    x += (double)f.x;
    y += (double)f.y;
    z += (double)f.z;
}

// Element k (runtime) of a small local array by selects on register values.
// The empty asm statements pin each element in registers first; otherwise
// the selects fold into an indexed load and the array moves to scratch.
template <typename T> __device__ inline v3<T> in_regs(v3<T> v) {
    asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z));
    return v;
}
template <typename T> __device__ inline v3<T> pick3(const v3<T> a[3], int k) {
    const v3<T> a0 = in_regs(a[0]), a1 = in_regs(a[1]), a2 = in_regs(a[2]);
    return k == 0 ? a0 : (k == 1 ? a1 : a2);
}
template <typename T> __device__ inline v3<T> pick4(const v3<T> a[4], int k) {
    const v3<T> a0 = in_regs(a[0]), a1 = in_regs(a[1]), a2 = in_regs(a[2]), a3 = in_regs(a[3]);
    return k == 0 ? a0 : (k == 1 ? a1 : (k == 2 ? a2 : a3));
}
template <int LPB> __device__ inline double lane_sum(double v) {
    for (int o = LPB / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Bonded terms of one block of beads [s0, s0 + NB), each quad, angle and
// bond computed once: quads starting at s0-3 .. s0+NB-1, angles and bonds
// starting at s0-2 .. s0+NB-1 (the few starting left of the block are
// computed again by the block on the left: the same inputs give the same
// bits).  Forces per bead position and energies per term, in FP64.
constexpr int QUAD_TERMS[4] = {T_CHIRAL, T_CND, T_SND, T_HEUR_DIH};
constexpr int ANG_TERMS[2] = {T_NAT_ANG, T_HEUR_ANG};
// The force kernel's workgroup: two wavefronts, one for the block's quads
// and one for its angles and bonds, so that no wavefront of the block waits
// idle for the bonded phase.
constexpr int FBLK = 128;
// Partners of a bead are mostly on its own chain, near it in sequence (on
// evolved 10k-bead boxes of 250-bead chains, 77 % of PID partners lie
// within 64 beads): each block stages the positions and residues of the
// beads within WIN_H (plus one) of its own in LDS, and reads a partner and
// its chain neighbours from there when they fall inside, from global
// memory otherwise -- the scattered gathers are what occupy the texture
// path.
constexpr int WIN_H = 64;
template <int LPB> struct BondedLds {
    static constexpr int NB = FBLK / LPB, NQ = NB + 3, NA = NB + 2;
    static constexpr int NW = NB + 2 * WIN_H + 2;  // window beads [s0 - WIN_H - 1, s0 + NB + WIN_H + 1)
    double2 wxy[NW];
    double wz[NW];
    int8_t wat[NW];
    v3<double> qF[NQ][4];
    double qV[NQ][4];
    v3<double> aF[NA][3];
    double aV[NA][2];
    v3<double> tF[NA][2];
    double tV[NA];
};

// Quad of beads first..first+3: chirality, native dihedrals, heuristic
// dihedral (xmd/src/forces/chiral.cpp:15-37, dihedral/complex_native.cpp:
// 13-49, simple_native.cpp:13-45, dihedral/heurestic.cpp:29-70).
template <typename T, int LPB>
__device__ inline void quad_task(const Dev &D, const TermParams<T> &P, uint32_t mask, const double *Pp, int first,
                                 BondedLds<LPB> &L, int q) {
    v3<double> F[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double V[4] = {0, 0, 0, 0};
    const int n = D.nslot;
    const bool ok = first >= 0 && first + 3 < n && D.chain[first] >= 0 && D.chain[first] == D.chain[first + 3] &&
                    !D.err[D.slot_rep[first]];
    if (ok && (mask & ((1u << T_CHIRAL) | (1u << T_CND) | (1u << T_SND) | (1u << T_HEUR_DIH)))) {
        const v3<double> q0 = ldp(Pp, first), q1 = ldp(Pp, first + 1), q2 = ldp(Pp, first + 2),
                         q3 = ldp(Pp, first + 3);
        const DihFrame<T> fr = dih_frame<T>(q1 - q0, q2 - q1, q3 - q2);
        v3<T> out[4];
        auto acc = [&](int term, T v) {
            for (int k = 0; k < 4; ++k) F[k] = F[k] + cvt<double>(out[k]);
            V[term] += (double)v;
        };
        if (mask & (1u << T_CHIRAL)) {
            const double nc = D.chir_nat[first];
            if (nc == nc) acc(0, chiral<T>(fr.r12, fr.r23, fr.r34, P.e_chi, (T)nc, (T)D.chir_fac[first], out));
        }
        const double ph = D.phi_nat[first];
        if (ph == ph) {
            if (mask & (1u << T_CND)) acc(1, nat_dihedral<T>(fr, 0, P.CDA, P.CDB, P.CDH, (T)ph, out));
            if (mask & (1u << T_SND)) acc(2, nat_dihedral<T>(fr, 1, P.CDA, P.CDB, P.CDH, (T)ph, out));
        } else if ((mask & (1u << T_HEUR_DIH)) && (D.bflags[first] & B_HDIH)) {
            // typed by the residues of beads 2 and 3
            const double *cd = D.heur_dih_c + 6 * (3 * heur_class(D.atype[first + 1]) + heur_class(D.atype[first + 2]));
            T c[6];
            for (int k = 0; k < 6; ++k) c[k] = (T)cd[k];
            acc(3, heur_dihedral<T>(fr, c, out));
        }
    }
    for (int k = 0; k < 4; ++k) {
        L.qF[q][k] = F[k];
        L.qV[q][k] = V[k];
    }
}

// Angle of beads first..first+2 (native, else heuristic; angle/native.cpp:
// 13-39, angle/heurestic.cpp:29-59) and the bond (first, first+1)
// (tether.cpp:21-36, FP64 in both forms: stiff bonds cancel in r - r_nat).
template <typename T, int LPB>
__device__ inline void angle_bond_task(const Dev &D, const Derived &G, const TermParams<T> &P, uint32_t mask,
                                       const double *Pp, int first, BondedLds<LPB> &L, int a) {
    v3<double> F[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, B[2] = {{0, 0, 0}, {0, 0, 0}};
    double V[2] = {0, 0}, Vb = 0;
    const int n = D.nslot;
    const int ch = first >= 0 && first < n ? D.chain[first] : -1;
    const bool live = ch >= 0 && !D.err[D.slot_rep[first]];
    if (live && first + 2 < n && D.chain[first + 2] == ch && (mask & ((1u << T_NAT_ANG) | (1u << T_HEUR_ANG)))) {
        const v3<double> q0 = ldp(Pp, first), q1 = ldp(Pp, first + 1), q2 = ldp(Pp, first + 2);
        const double th = D.theta_nat[first];
        if (th == th) {
            if (mask & (1u << T_NAT_ANG)) {
                v3<T> out[3];
                V[0] = (double)nat_angle<T>(q1 - q0, q2 - q1, P.ang_k, (T)th, out);
                for (int k = 0; k < 3; ++k) F[k] = cvt<double>(out[k]);
            }
        } else if ((mask & (1u << T_HEUR_ANG)) && (D.bflags[first] & B_HANG)) {
            // typed by the residues of beads 1 and 2, FP64 in both forms
            const double *ca = D.heur_ang_c + 7 * (3 * heur_class(D.atype[first]) + heur_class(D.atype[first + 1]));
            V[1] = heur_angle(q1 - q0, q2 - q1, ca, F);
        }
    }
    if (live && first + 1 < n && D.chain[first + 1] == ch && (mask & (1u << T_TETHER)))
        Vb = tether<double>(ldp(Pp, first + 1) - ldp(Pp, first), G.H1, G.H2, D.tether_nat[first], B);
    for (int k = 0; k < 3; ++k) L.aF[a][k] = F[k];
    L.aV[a][0] = V[0];
    L.aV[a][1] = V[1];
    L.tF[a][0] = B[0];
    L.tF[a][1] = B[1];
    L.tV[a] = Vb;
}

constexpr int SS_TYPES = 400;  // residue pairs 20 a1 + a2

// Whole wavefronts per task kind (quads from wave 0, angles and bonds from
// the next free wave) so that no wavefront runs both code paths.
template <typename T, int LPB>
__device__ inline void bonded_phase(const Dev &D, const Derived &G, const TermParams<T> &P, uint32_t mask,
                                    const double *Pp, int s0, BondedLds<LPB> &L) {
    using BL = BondedLds<LPB>;
    constexpr int A0 = (BL::NQ + WAVE - 1) / WAVE * WAVE;
    for (int task = threadIdx.x; task < A0 + BL::NA; task += FBLK) {
        if (task < BL::NQ) quad_task<T, LPB>(D, P, mask, Pp, s0 - 3 + task, L, task);
        else if (task >= A0) angle_bond_task<T, LPB>(D, G, P, mask, Pp, s0 - 2 + (task - A0), L, task - A0);
    }
    __syncthreads();
}

template <typename T, bool EVAL, int LPB>
__global__ __launch_bounds__(FBLK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
void k_force(Dev D, Derived G, TermTab<T> TT, int p, int mutate) {
    const int n = D.nslot;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int t = blk * FBLK + threadIdx.x;
    const int s = t / LPB, sub = t % LPB;
    const int s0 = blk * (FBLK / LPB);  // the block's first bead
    __shared__ BondedLds<LPB> L;
    // the PID bundle's ss LJ table (depth, r_min, r_max by residue pair;
    // float copies in the FP32 form), read per bundle from LDS
    using T4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    __shared__ T4 ss_tab[SS_TYPES];
    {
        const T4 *src = reinterpret_cast<const T4 *>(sizeof(T) == 4 ? (const void *)D.ss_pack_f : (const void *)D.ss_pack);
        for (int q = threadIdx.x; q < SS_TYPES; q += FBLK) ss_tab[q] = src[q];
    }
    // the window (slots clamped to the slot range: a partner is always in
    // the bead's own replica, so a clamped slot is never read as one)
    const int w0 = s0 - WIN_H - 1;
    for (int q = threadIdx.x; q < BondedLds<LPB>::NW; q += FBLK) {
        const int slot = min(max(w0 + q, 0), D.nslot - 1);
        const double *a = D.pos[p] + 4 * (size_t)slot;
        L.wxy[q] = *reinterpret_cast<const double2 *>(a);
        L.wz[q] = a[2];
        L.wat[q] = D.atype[slot];
    }
    bonded_phase<T, LPB>(D, G, TT.P, TT.mask, D.pos[p], s0, L);  // ends with a barrier
    // a wavefront's beads lie in one 64-aligned slot range, hence one
    // replica: its index (and so the box) is wave-uniform, in scalar registers
    const int rep = __builtin_amdgcn_readfirstlane(D.slot_rep[s]);
    const uint8_t bf = D.bflags[s];
    // a halted replica (error word set) is skipped: nothing it computes
    // would be used, and its wall pair states must not change
    const bool active = (bf & B_ACTIVE) && !D.err[rep];
    const double *Pp = D.pos[p];
    const TermParams<T> &P = TT.P;
    const uint32_t mask = TT.mask;
    double Fx = 0, Fy = 0, Fz = 0, V = 0;
    double Vt[EVAL ? NUM_TERMS : 1];
    for (int k = 0; k < (EVAL ? NUM_TERMS : 1); ++k) Vt[k] = 0;
    auto addV = [&](int term, double v) {
        V += v;
        if constexpr (EVAL) Vt[term] += v;
    };
    double tpx = 0, tpy = 0, tpz = 0, tnx = 0, tny = 0, tnz = 0;  // PID spill to s-1, s+1
    if (active) {
        const v3<double> ri = ldp(Pp, s);
        const Box box = D.box[rep];
        // bonded terms: the block computed each quad, angle and tether once
        // (bonded_phase); the bead sums the ones it belongs to, in a fixed
        // order: quads starting at s-3..s, angles at s-2..s, bonds (s-1, s)
        // and (s, s+1); an energy is the first bead's
        const int lo = s - s0;
        for (int c = sub; c < 9; c += LPB) {
            if (c < 4) {
                const int q = lo + c;  // quad starting at s - 3 + c, this bead at position 3 - c
                add3(Fx, Fy, Fz, L.qF[q][3 - c]);
                if (c == 3)
                    for (int k = 0; k < 4; ++k) addV(QUAD_TERMS[k], L.qV[q][k]);
            } else if (c < 7) {
                const int a = lo + (c - 4);  // angle starting at s - 2 + (c - 4)
                add3(Fx, Fy, Fz, L.aF[a][6 - c]);
                if (c == 6)
                    for (int k = 0; k < 2; ++k) addV(ANG_TERMS[k], L.aV[a][k]);
            } else if (c == 7) {
                add3(Fx, Fy, Fz, L.tF[lo + 1][1]);  // bond (s - 1, s)
            } else {
                add3(Fx, Fy, Fz, L.tF[lo + 2][0]);  // bond (s, s + 1)
                addV(T_TETHER, L.tV[lo + 2]);
            }
        }
        // native contacts (Go)
        if (mask & (1u << T_GO)) {
            for (int k = D.nat_start[s] + sub; k < D.nat_start[s + 1]; k += LPB) {
                if (!D.nat_act[k]) continue;
                int j = D.nat_j[k];
                v3<double> d = wrap(box, ldp(Pp, j) - ri);
                FV<T> f = go_pair<T>(cvt<T>(d), P.go_depth, P.go_d12, (T)D.nat_d[k]);
                add3(Fx, Fy, Fz, f.F);
                if (s < j) addV(T_GO, (double)f.V);
            }
        }
        // disulfide bonds (native_ss.cpp:27-40), FP64 in both forms
        if (mask & (1u << T_NAT_SS)) {
            for (int k = D.ss_start[s] + sub; k < D.ss_start[s + 1]; k += LPB) {
                if (!D.ss_act[k]) continue;
                const int j = D.ss_j[k];
                const v3<double> d = wrap(box, ldp(Pp, j) - ri);
                const double r = norm(d);
                const VdV<double> e = harmonic<double>(G.ss_H1, 0.0, G.ss_equil, r);
                add3(Fx, Fy, Fz, d * (e.dV / r));
                if (s < j) addV(T_NAT_SS, e.V);
            }
        }
        // neighbour row: Pauli, Debye-Hueckel, PID
        uint32_t rmask = mask & ((1u << T_PAULI) | (1u << T_CONST_ES) | (1u << T_REL_ES) | (1u << T_PID));
        if (rmask) {
            const int len = D.nl_len[s];
            const int base = D.slice_start[s >> 6] + (s & 63);
            const bool ces_on = mask & (1u << T_CONST_ES), res_on = mask & (1u << T_REL_ES);
            const T q_s = (T)D.charge[D.atype[s]];
            const v3<double> rsp = (bf & B_PREV) ? ldp(Pp, s - 1) : ri;
            const v3<double> rsn = (bf & B_NEXT) ? ldp(Pp, s + 1) : ri;
            // each lane loads its entries (partner and flags in one word)
            // two iterations ahead, so an iteration waits on one round of
            // position gathers only
            int e = base + 64 * sub;
            uint32_t ent_n = sub < len ? D.nl_e[e] : 0u;
            uint32_t ent_nn = sub + LPB < len ? D.nl_e[e + 64 * LPB] : 0u;
            for (int k = sub; k < len; k += LPB, e += 64 * LPB) {
                const uint8_t fl = nle_flag(ent_n);
                const int j = nle_j(ent_n);
                ent_n = ent_nn;
                if (k + 2 * LPB < len) ent_nn = D.nl_e[e + 128 * LPB];
                if (!(fl & (F_PAULI | F_ES | F_PID))) continue;
                // the partner and its chain neighbours (read for PID
                // entries, whose partners have both), and its residue
                v3<double> rj, rjp, rjn;
                int aj;
                const int w = j - w0;
                if (w >= 1 && w + 1 < BondedLds<LPB>::NW) {
                    const double2 a = L.wxy[w], b = L.wxy[w - 1], c = L.wxy[w + 1];
                    rj = {a.x, a.y, L.wz[w]};
                    rjp = {b.x, b.y, L.wz[w - 1]};
                    rjn = {c.x, c.y, L.wz[w + 1]};
                    aj = L.wat[w];
                } else {
                    rj = ldp(Pp, j);
                    rjp = (fl & F_PID) ? ldp(Pp, j - 1) : rj;
                    rjn = (fl & F_PID) ? ldp(Pp, j + 1) : rj;
                    aj = D.atype[j];
                }
                const v3<double> dd = wrap(box, rj - ri);
                const v3<T> d = cvt<T>(dd);
                const bool owner = s < j;
                if ((fl & F_PAULI) && (mask & (1u << T_PAULI))) {
                    FV<T> f = pauli_pair<T>(d, P.pauli_depth, P.pauli_d12, P.pauli_r_excl);
                    add3(Fx, Fy, Fz, f.F);
                    if (owner) addV(T_PAULI, (double)f.V);
                }
                if ((fl & F_ES) && (ces_on || res_on)) {
                    // both variants run over the one list when both are on (D31)
                    T q1q2 = owner ? q_s * (T)D.charge[aj] : (T)D.charge[aj] * q_s;
                    if (ces_on) {
                        FV<T> f = es_pair<T>(d, q1q2, P.es_Vf, P.es_sdi, 0);
                        add3(Fx, Fy, Fz, f.F);
                        if (owner) addV(T_CONST_ES, (double)f.V);
                    }
                    if (res_on) {
                        FV<T> f = es_pair<T>(d, q1q2, P.res_Vf, P.res_sdi, 1);
                        add3(Fx, Fy, Fz, f.F);
                        if (owner) addV(T_REL_ES, (double)f.V);
                    }
                }
                if ((fl & F_PID) && (mask & (1u << T_PID))) {
                    // The bundle from this bead's side: s in the first slot.
                    // lambda(psi_1) lambda(psi_2) LJ(r) is symmetric in the
                    // two sides, so out[0..2] are the reference's forces on
                    // (s-1, s, s+1) whichever bead of the pair is i1; the
                    // energy is counted by the owner (s = i1).
                    const int a_s = D.atype[s];
                    const int ty = owner ? 20 * a_s + aj : 20 * aj + a_s;
                    v3<T> out[6];
                    const T4 sst = ss_tab[ty];
                    T v = pid_bundle<T>(P, sst.x, sst.y, sst.z,
                                        pid_side<T>(ri - rsp, rsn - ri, rsn - rsp),
                                        pid_side<T>(rj - rjp, rjn - rj, rjn - rjp), dd, out);
                    add3(Fx, Fy, Fz, out[1]);
                    add3(tpx, tpy, tpz, out[0]);
                    add3(tnx, tny, tnz, out[2]);
                    if (owner) addV(T_PID, (double)v);
                }
            }
        }
        // walls and AFM tips (walls/solid.cpp:19-30, walls/lj_attr.cpp:7-49,
        // velocity_afm.cpp:12-24, force_afm.cpp:11-14; D25-D27), FP64 in
        // both forms, on the bead's last lane
        constexpr uint32_t WALL_AFM =
            (1u << T_SOLID_WALL) | (1u << T_LJ_WALL) | (1u << T_VEL_AFM) | (1u << T_FORCE_AFM);
        if ((mask & WALL_AFM) && sub == LPB - 1) {
            // ISO evaluation here: the pair states test saturation == 0 and
            // == 1 exactly (lj_attr.cpp:36-43), and a fused 1 - m c^-1
            // misses both
#pragma clang fp contract(off)
            const double tnow = D.tt[p][rep];
            const v3<double> rw = wrap(box, ri);  // box.wrap of the position
            if (mask & (1u << T_SOLID_WALL))
                for (int w = 0; w < G.n_solid; ++w) {
                    const double *pl = D.solid_w + 6 * w;
                    const v3<double> nn = {pl[3], pl[4], pl[5]};
                    const double x = dot(nn, rw - v3<double>{pl[0], pl[1], pl[2]});
                    if (x < G.solid_cut) {
                        const double x3 = x * x * x;
                        const double Vw = G.solid_eps / (9.0 * (x3 * x3 * x3));
                        add3(Fx, Fy, Fz, nn * (9.0 * Vw / x));
                        addV(T_SOLID_WALL, Vw);
                    }
                }
            if (mask & (1u << T_LJ_WALL))
                for (int w = 0; w < G.n_ljw; ++w) {
                    const size_t k = (size_t)s * G.n_ljw + w;
                    const double *pl = D.lj_w + 6 * w;
                    const v3<double> nn = {pl[3], pl[4], pl[5]};
                    const int st = D.ljw_status[k];
                    if (st == 0) {
                        const double x = dot(nn, rw - v3<double>{pl[0], pl[1], pl[2]});
                        if (fabs(x) < G.ljw_min && mutate) {
                            D.ljw_status[k] = 1;
                            const v3<double> j = rw - nn * x;
                            D.ljw_joint[3 * k] = j.x, D.ljw_joint[3 * k + 1] = j.y, D.ljw_joint[3 * k + 2] = j.z;
                            D.ljw_ref[k] = tnow;
                        }
                        continue;
                    }
                    double sat = fmin(tnow - D.ljw_ref[k], G.ljw_cyc) * G.ljw_cyc_inv;
                    if (st == 2) sat = 1.0 - sat;
                    const v3<double> j = {D.ljw_joint[3 * k], D.ljw_joint[3 * k + 1], D.ljw_joint[3 * k + 2]};
                    const v3<double> r12 = j - rw;
                    const double rn = 1.0 / norm(r12);
                    const VdV<double> e = lj<double>(1.0, G.ljw_min, rn);
                    addV(T_LJ_WALL, e.V * sat);
                    add3(Fx, Fy, Fz, r12 * (sat * e.dV * rn));
                    if (mutate) {
                        if (st == 1 && sat == 1.0) {
                            if (G.ljw_factor * G.ljw_min * rn < 1.0) {
                                D.ljw_status[k] = 2;
                                D.ljw_ref[k] = tnow;
                            }
                        } else if (st == 2 && sat == 0.0) {
                            D.ljw_status[k] = 0;
                        }
                    }
                }
            if (mask & (1u << T_VEL_AFM))
                for (int k = D.vafm_start[s]; k < D.vafm_start[s + 1]; ++k) {
                    const v3<double> tip = v3<double>{D.vafm_orig[3 * k], D.vafm_orig[3 * k + 1], D.vafm_orig[3 * k + 2]} +
                                           tnow * v3<double>{D.vafm_v[3 * k], D.vafm_v[3 * k + 1], D.vafm_v[3 * k + 2]};
                    const v3<double> ra = tip - ri;
                    const double rn = norm(ra);
                    if (rn == 0.0) continue;  // the tip on its bead: no force (D27)
                    const VdV<double> e = harmonic<double>(G.vafm_H1, G.vafm_H2, 0.0, rn);
                    add3(Fx, Fy, Fz, ra * (e.dV / rn));
                }
            if (mask & (1u << T_FORCE_AFM))
                for (int k = D.fafm_start[s]; k < D.fafm_start[s + 1]; ++k) {
                    Fx += D.fafm_F[3 * k];
                    Fy += D.fafm_F[3 * k + 1];
                    Fz += D.fafm_F[3 * k + 2];
                }
        }
        // quasi-adiabatic contacts through the bead's contact slots
        if (mask & (1u << T_QA)) {
            const int q = D.qn[s], tot = D.ct_cap[D.nrep];
            for (int k = sub; k < q; k += LPB) {
                int ci = D.qslot[k * n + s];
                double fx = D.ct_F[ci], fy = D.ct_F[tot + ci], fz = D.ct_F[2 * tot + ci];
                if (D.ct_i1[ci] == s) {
                    Fx += fx;
                    Fy += fy;
                    Fz += fz;
                    addV(T_QA, D.ct_V[ci]);
                } else {
                    Fx -= fx;
                    Fy -= fy;
                    Fz -= fz;
                }
            }
        }
    }
    Fx = lane_sum<LPB>(Fx);
    Fy = lane_sum<LPB>(Fy);
    Fz = lane_sum<LPB>(Fz);
    V = lane_sum<LPB>(V);
    tpx = lane_sum<LPB>(tpx);
    tpy = lane_sum<LPB>(tpy);
    tpz = lane_sum<LPB>(tpz);
    tnx = lane_sum<LPB>(tnx);
    tny = lane_sum<LPB>(tny);
    tnz = lane_sum<LPB>(tnz);
    if constexpr (EVAL)
        for (int k = 0; k < NUM_TERMS; ++k) Vt[k] = lane_sum<LPB>(Vt[k]);
    if (active && sub == 0) {
        if constexpr (EVAL)
            for (int k = 0; k < NUM_TERMS; ++k) D.Vterm[k * n + s] = Vt[k];
        D.Fbuf[s] = Fx;
        D.Fbuf[n + s] = Fy;
        D.Fbuf[2 * n + s] = Fz;
        D.Vbead[s] = V;
        double *ps = D.pid_side;
        ps[s] = tpx;
        ps[n + s] = tpy;
        ps[2 * n + s] = tpz;
        ps[3 * n + s] = tnx;
        ps[4 * n + s] = tny;
        ps[5 * n + s] = tnz;
    }
}

}  // namespace xmd_k
