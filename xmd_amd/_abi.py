"""ctypes mirror of include/xmd_hip.h (structs, enums, entry points).

The C header is the contract; this module only restates its layout for the
Python host side.  tests/test_abi.py checks every field offset against the
header compiled with the system C compiler.
"""
import ctypes as C
import os

NUM_AA = 20
NUM_SS = NUM_AA * NUM_AA
QA_NUM_TYPES = 4 + NUM_SS

# xmd_status
OK, ERR_ARG, ERR_HIP, ERR_CAPACITY, ERR_NONFINITE, ERR_STATE, ERR_NO_DEVICE = range(7)

# xmd_term
TERMS = ["tether", "chiral", "nat_ang", "cnd", "snd", "pauli", "go",
         "const_es", "rel_es", "qa", "pid", "nat_ss", "solid_wall", "lj_wall", "vel_afm", "force_afm",
         "heur_ang", "heur_dih"]
TERM = {name: k for k, name in enumerate(TERMS)}
NUM_TERMS = len(TERMS)
ALL_TERMS = (1 << NUM_TERMS) - 1

# xmd_list_id
LISTS = ["nl_all", "nl_native", "nl_non_native", "go", "pauli", "es", "pid", "qa_free", "nl_nat_ss", "nat_ss"]
LIST = {name: k for k, name in enumerate(LISTS)}

PREC_FP64, PREC_MIXED = 0, 1
LAMBDA_COSINE, LAMBDA_ALGEBRAIC = 0, 1

_d, _i32, _i8 = C.c_double, C.c_int32, C.c_int8


class XmdParams(C.Structure):
    _fields_ = [
        ("dt", _d), ("gamma_factor", _d),
        ("langevin_enabled", _i32), ("precision", _i32), ("rng_exact", _i32), ("stream_groups", _i32),
        ("pad_factor", _d), ("max_cutoff_override", _d),
        ("tether_enabled", _i32), ("_pad1", _i32),
        ("tether_H1", _d), ("tether_H2", _d), ("tether_def_length", _d),
        ("chir_enabled", _i32), ("_pad2", _i32), ("chir_e", _d),
        ("nat_ang_enabled", _i32), ("_pad3", _i32), ("nat_ang_k", _d),
        ("cnd_enabled", _i32), ("snd_enabled", _i32),
        ("cnd_CDA", _d), ("cnd_CDB", _d), ("snd_CDH", _d),
        ("pauli_enabled", _i32), ("_pad4", _i32), ("pauli_r_excl", _d), ("pauli_depth", _d),
        ("go_enabled", _i32), ("_pad5", _i32), ("go_depth", _d),
        ("const_es_enabled", _i32), ("rel_es_enabled", _i32),
        ("const_es_screen", _d), ("const_es_permittivity", _d),
        ("rel_es_screen", _d), ("rel_es_factor", _d),
        ("aa_charge", _d * NUM_AA),
        ("lj_bb_rmin", _d), ("lj_bb_depth", _d), ("lj_bs_rmin", _d), ("lj_bs_depth", _d),
        ("ss_rmax", _d * NUM_SS), ("ss_depth", _d * NUM_SS),
        ("ss_sinking", _i32), ("_pad6", _i32),
        ("pid_enabled", _i32), ("pid_lambda_version", _i32),
        ("pid_bbp_alpha", _d), ("pid_bbp_psi0", _d), ("pid_bbp_rmin", _d), ("pid_bbp_depth", _d),
        ("pid_bbm_alpha", _d), ("pid_bbm_psi0", _d), ("pid_bbm_rmin", _d), ("pid_bbm_depth", _d),
        ("pid_ss_alpha", _d), ("pid_ss_psi0", _d),
        ("qa_enabled", _i32), ("_pad7", _i32),
        ("qa_phase_duration", _d), ("qa_breaking_factor", _d),
        ("qa_min_cos_hr", _d), ("qa_min_cos_hh", _d), ("qa_max_cos_nr", _d),
        ("aa_limits", (_i8 * 4) * NUM_AA),
        ("aa_polarization", _i8 * NUM_AA),
        ("force_lanes", _i32),
        ("nat_ss_enabled", _i32), ("nat_ss_H1", _d), ("nat_ss_equil_dist", _d),
        ("solid_walls_enabled", _i32), ("lj_walls_enabled", _i32),
        ("solid_walls_depth", _d), ("solid_walls_cutoff", _d),
        ("lj_walls_min_dist", _d), ("lj_walls_breaking_factor", _d), ("lj_walls_phase_duration", _d),
        ("vel_afm_enabled", _i32), ("force_afm_enabled", _i32),
        ("vel_afm_H1", _d), ("vel_afm_H2", _d),
        ("heur_ang_enabled", _i32), ("heur_dih_enabled", _i32),
        ("heur_ang_coeffs", (_d * 7) * 9), ("heur_dih_coeffs", (_d * 6) * 9),
        ("row_width", _i32), ("fp_capacity", _i32), ("ct_capacity", _i32),
    ]


class XmdTopology(C.Structure):
    _fields_ = [
        ("n_replicas", _i32), ("n_beads", _i32), ("n_chains", _i32), ("n_native", _i32),
        ("replica_offset", C.POINTER(_i32)), ("chain_offset", C.POINTER(_i32)),
        ("atype", C.POINTER(_i8)), ("r0", C.POINTER(_d)), ("box", C.POINTER(_d)),
        ("temperature", C.POINTER(_d)), ("rng_state", C.POINTER(C.c_uint64)),
        ("nat_i1", C.POINTER(_i32)), ("nat_i2", C.POINTER(_i32)), ("nat_dist", C.POINTER(_d)),
        ("tether_len", C.POINTER(_d)), ("nat_theta", C.POINTER(_d)), ("nat_phi", C.POINTER(_d)),
        ("nat_chir", C.POINTER(_d)), ("nat_chir_factor", C.POINTER(_d)),
        ("n_ssbonds", _i32), ("ss_i1", C.POINTER(_i32)), ("ss_i2", C.POINTER(_i32)),
        ("n_solid_walls", _i32), ("solid_walls", C.POINTER(_d)),
        ("n_lj_walls", _i32), ("lj_walls", C.POINTER(_d)),
        ("n_vel_afm", _i32), ("vel_afm_idx", C.POINTER(_i32)), ("vel_afm_orig", C.POINTER(_d)),
        ("vel_afm_v", C.POINTER(_d)),
        ("n_force_afm", _i32), ("force_afm_idx", C.POINTER(_i32)), ("force_afm_F", C.POINTER(_d)),
        ("heur_bonded", C.POINTER(C.c_uint8)),
    ]


class XmdQaContact(C.Structure):
    _fields_ = [
        ("i1", _i32), ("i2", _i32), ("type", C.c_int16), ("status", _i8), ("_pad", _i8),
        ("diff1", _i8 * 4), ("diff2", _i8 * 4), ("ref_time", _d),
    ]


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array (None stays NULL)."""
    if arr is None:
        return C.cast(None, C.POINTER(ctype))
    return arr.ctypes.data_as(C.POINTER(ctype))


def repo_root():
    return os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
