// common.hpp -- declarations shared by the engine's two translation units
// (engine.hip and forces.hip): constants, the derived parameters, the table
// of device pointers, slot-data access helpers and the force launch.
#pragma once
#include <hip/hip_runtime.h>

#include "xmd_dev.hpp"

namespace xmd_k {
using namespace xmd_dev;

constexpr int WAVE = 64;
constexpr int BLK = 256;
constexpr int MAXQ = 16;
constexpr int NUM_TERMS = 18;

// F_NATSS marks a disulfide pair of the list (native_ssbonds): F_NATIVE is
// set too, so no non-native term takes it.
enum : uint8_t { F_NATIVE = 1, F_PAULI = 2, F_ES = 4, F_PID = 8, F_QAFREE = 16, F_NATSS = 32 };
// A neighbour-row entry packs the partner slot and the flags into one word
// (one load per entry); slots are below 2^NLE_SHIFT.
constexpr int NLE_SHIFT = 26;
constexpr uint32_t NLE_JMASK = (1u << NLE_SHIFT) - 1u;
__host__ __device__ inline uint32_t nle_pack(int j, uint8_t f) { return (uint32_t)j | ((uint32_t)f << NLE_SHIFT); }
__host__ __device__ inline int nle_j(uint32_t x) { return (int)(x & NLE_JMASK); }
__host__ __device__ inline uint8_t nle_flag(uint32_t x) { return (uint8_t)(x >> NLE_SHIFT); }
// B_HANG / B_HDIH: the model's angle (dihedral) starting at this bead has no
// native value, so the heuristic term takes it when enabled.
enum : uint8_t { B_PREV = 1, B_NEXT = 2, B_ACTIVE = 4, B_HANG = 8, B_HDIH = 16 };
// E_HOLD is no error: the host holds the replica (every kernel leaves it as
// it is, like a halted one) while it steps another alone.
enum : int { E_NL_CAP = 1, E_FP_CAP = 2, E_CT_CAP = 4, E_SLOT = 8, E_NONFINITE = 16, E_ZERO_DRAW = 32, E_HOLD = 64 };
enum : int {
    T_TETHER = 0, T_CHIRAL, T_NAT_ANG, T_CND, T_SND, T_PAULI, T_GO, T_CONST_ES, T_REL_ES, T_QA, T_PID, T_NAT_SS,
    T_SOLID_WALL, T_LJ_WALL, T_VEL_AFM, T_FORCE_AFM, T_HEUR_ANG, T_HEUR_DIH
};

// Derived parameters in double (list and QA decisions, integrator).
struct Derived {
    double dt, dt_inv, gamma, half_dt2;
    double H1, H2;  // tether constants (FP64 in both forms)
    double pad, req_r;
    double pauli_min_inv, es_min_inv, pid_min_inv, qa_min_inv;
    int pauli_on, es_on, pid_on, qa_on, go_on;
    int lang_on;
    // disulfide bonds: harmonic H1 about the equilibrium distance, current
    // while within ss_cut + pad (FP64 in both forms)
    int ss_on;
    double ss_H1, ss_equil, ss_cut;
    // walls (planes in Dev) and AFM tips, FP64 in both forms
    int n_solid, n_ljw;
    double solid_eps, solid_cut, ljw_min, ljw_factor, ljw_cyc, ljw_cyc_inv;
    double vafm_H1, vafm_H2;
    QaParams qa;
};

// Every device pointer of an engine (passed by value to kernels).
struct Dev {
    int nslot, nrep;
    // replicas
    const int *rep_slot_off, *rep_n;
    const Box *box;
    const double *noise_sd;
    uint64_t *rng[2];
    double *tt[2];
    // largest displacement since the list build of each 64-slot slice (bit
    // pattern of a non-negative double, NaN above all), per parity
    unsigned long long *slice_maxd[2];
    int *rb_now, *force_rb, *any_rb, *err;
    long long *n_rebuild;
    long long *n_done;  // steps each replica has taken (halted ones stop counting)
    int *need_w;        // longest row a rebuild of the replica wanted beyond the row width
    int *need_fp;       // QA free pairs a rebuild of the replica found beyond its capacity
    const int *cell_off, *cell_dim;  // [nrep+1], [3*nrep]
    int *cell_cnt, *cell_start, *cell_end, *cell_cursor, *cell_beads, *bead_cell;
    // slots
    const int *slot_rep, *chain;
    const uint8_t *bflags;
    const int8_t *atype;
    double *pos[2];
    double *y;  // y1..y5, 15*nslot
    double *orig_r, *Fbuf, *pid_side, *Vbead, *Vterm;
    const double *tether_nat, *theta_nat, *phi_nat, *chir_nat, *chir_fac;
    // native contacts (rows of slot s: nat_start[s]..nat_start[s+1])
    const int *nat_start, *nat_j;
    const double *nat_d;
    uint8_t *nat_act;
    // disulfide bonds (rows of slot s: ss_start[s]..ss_start[s+1])
    const int *ss_start, *ss_j;
    uint8_t *ss_act;
    // wall planes (6 doubles: origin, normal); LJ attractive wall state per
    // slot x wall (0 FREE, 1 FORMING_OR_FORMED, 2 BREAKING), joint, time
    const double *solid_w, *lj_w;
    int8_t *ljw_status;
    double *ljw_joint, *ljw_ref;
    // AFM tips of slot s: vafm_start[s].. (origin, velocity), fafm_start[s].. (force)
    const int *vafm_start, *fafm_start;
    const double *vafm_orig, *vafm_v, *fafm_F;
    // neighbour rows
    int nl_w;  // row capacity per slot (rows are fixed-width sliced ELL)
    // entry: partner slot in the low NLE_SHIFT bits, term flags above
    int *slice_start, *nl_len;
    uint32_t *nl_e;
    // QA free pairs, candidates
    const int *fp_base, *fp_cap, *fp_chunk_rep;
    int *fp_count, *fp_i1, *fp_i2, *fp_slice_cnt;
    int *fp_n;  // QA free pairs of each slot's row (set at rebuild)
    uint8_t *fp_vac;
    int *cand_cnt, *cand_fp, *cand_type;
    uint32_t *cand_d1, *cand_d2;
    // QA contacts, removals
    const int *ct_base, *ct_cap, *ct_chunk_rep;
    int *ct_count, *ct_i1, *ct_i2, *ct_type, *ct_status;
    double *ct_ref, *ct_F, *ct_V;
    uint32_t *ct_d1, *ct_d2;
    uint8_t *ct_vac;
    int *rm_cnt, *rm_idx;
    uint32_t *sync;
    int *qslot, *qn;
    // tables
    const double *ss_depth, *ss_rmin, *ss_rmax, *req_min_dist, *charge;
    // the PID ss LJ table packed per residue pair (depth, r_min, r_max, 0):
    // one 16- (float, for the FP32 form) or 32-byte (double) load per pair type
    const float *ss_pack_f;
    const double *ss_pack;
    const int8_t *ptype;
    const uint8_t *bm_fix;  // glibc correction table and its thresholds (bead_noise)
    const float *bm_thr;
    // heuristic bonded tables by pair type 3 c1 + c2 (c: 0 GLY, 1 PRO, 2
    // other): 7 angle polynomial coefficients, then 6 dihedral coefficients
    const double *heur_ang_c, *heur_dih_c;
    // rows of J^(256 2^j) (jrow), J^(256 w) for w < jwave_n (jwave) and
    // J^(4N) per replica (jadv); lane nibble table jnib
    const uint64_t *jrow, *jwave, *jnib, *jadv;
    int jwave_n;
};

__device__ inline v3<double> ld3(const double *a, int n, int s) { return {a[s], a[n + s], a[2 * n + s]}; }
// Positions and the positions at the last list build are AoS, 32 bytes per
// slot (x, y, z, 0): a partner gather reads one cache line, as a 16-byte
// and an 8-byte load.
__device__ inline v3<double> ldp(const double *a, int s) {
    const double2 xy = *reinterpret_cast<const double2 *>(a + 4 * (size_t)s);
    return {xy.x, xy.y, a[4 * (size_t)s + 2]};
}
__device__ inline void stp(double *a, int s, v3<double> v) {
    double2 *q = reinterpret_cast<double2 *>(a + 4 * (size_t)s);
    q[0] = make_double2(v.x, v.y);
    q[1] = make_double2(v.z, 0.0);
}
__device__ inline void st3(double *a, int n, int s, v3<double> v) {
    a[s] = v.x;
    a[n + s] = v.y;
    a[2 * n + s] = v.z;
}
__device__ inline int lane_id() { return threadIdx.x & (WAVE - 1); }

// Workgroups are dispatched to the 8 XCDs round-robin by block index, and
// each XCD has its own L2.  For a grid of nb blocks (a multiple of 8; other
// grids keep their order) this maps block b to a position so that XCD x
// covers the contiguous eighth
// [x nb/8, (x+1) nb/8) of the grid: the partner gathers of one XCD then
// stay within a few replicas, whose positions fit its L2.
constexpr int NUM_XCD = 8;
__device__ inline int xcd_block(int b, int nb) {
    return nb % NUM_XCD ? b : (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD;
}

// Parameters of the force kernel in its precision T, and the term mask.
template <typename T> struct TermTab {
    TermParams<T> P;
    uint32_t mask;
};

// Lanes per bead of the force kernel, and its launch (forces.hip, a
// translation unit of its own compiled with multiply-add contraction).
// Lanes per bead: 4, 8 (default) or 16, chosen by the autotuner.
constexpr int FORCE_LPB = 8;
// mutate: the launch is a step's force pass (LJ attractive wall pairs change
// state); 0 for evaluations.
void launch_force(bool mixed, bool eval, int lpb, int nslot, hipStream_t stream, const Dev &D, const Derived &G,
                  const TermTab<float> &Tf, const TermTab<double> &Td, int p, int mutate);

}  // namespace xmd_k
