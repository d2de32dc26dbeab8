"""ctypes mirror of include/mam4x.h.

The structures below must match the C declarations field for field; the
``test_abi`` CPU tests check sizes against the header's layout rules.
"""
import ctypes
import os

ABI_VERSION = 11

NSTATE = 52
NMODE = 4

OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_UNSUPPORTED = 3
ERR_NOMEM = 4

PREC_FP64 = 0
PREC_MIXED = 1

SCHED_AUTO = 0
SCHED_FUSED = 1
SCHED_QUEUE = 2

LAYOUT_LEV_COL = 0
LAYOUT_COL_LEV = 1

NPHASE = 3
PHASE_NAMES = ("setup", "soa", "finish")

PCNST = 40
PACK_STATE = 0
PACK_TENDENCY = 1

GAS_PCNST = 31
NBUDGET = 54  # MAM4X_NBUDGET: the 52 state fields, then the closures
BUDGET_SULFUR = 52
BUDGET_ORGANIC = 53
NPROCESS = 5
NOALLOC_MAX_CELLS = 524288  # MAM4X_NOALLOC_MAX_CELLS
PROC_CONDENSATION = 0
PROC_RENAMING = 1
PROC_NUCLEATION = 2
PROC_COAGULATION = 3
PROC_RENAMING_CLOUD_BORNE = 4

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class Config(ctypes.Structure):
    _fields_ = [
        ("do_cond", ctypes.c_int),
        ("do_rename", ctypes.c_int),
        ("do_newnuc", ctypes.c_int),
        ("do_coag", ctypes.c_int),
        ("gaexch_h2so4_uptake_optaa", ctypes.c_int),
        ("newnuc_h2so4_conc_optaa", ctypes.c_int),
        ("n_so4_monolayers_pcage", ctypes.c_uint),
        ("do_calcsize", ctypes.c_int),
        ("calcsize_do_aitacc_transfer", ctypes.c_int),
        ("calcsize_do_adjust", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("schedule", ctypes.c_int),
        ("soa_blocks_per_cu", ctypes.c_int),
        ("ntsubstep", ctypes.c_int),
        ("soa_chunk", ctypes.c_int),
        ("soa_short_chunk", ctypes.c_int),
        ("soa_short_bucket", ctypes.c_int),
        ("levels_per_wave", ctypes.c_int),
    ]


class Inputs(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("state", _vp * NSTATE),
        ("temperature", _vp),
        ("pressure", _vp),
        ("height", _vp),
        ("vapor_mixing_ratio", _vp),
        ("cloud_fraction", _vp),
        ("pblh", _vp),
        ("dgncur_a", _vp * NMODE),
        ("dgncur_awet", _vp * NMODE),
        ("wetdens", _vp * NMODE),
        ("pregaschem_h2so4", _vp),
        ("pregaschem_soag", _vp),
        ("precldchem", _vp * NSTATE),
        ("hydrostatic_dp", _vp),
    ]


class Tangent(ctypes.Structure):
    _fields_ = [
        ("state", _vp * NSTATE),
        ("temperature", _vp),
        ("pressure", _vp),
        ("vapor_mixing_ratio", _vp),
        ("tendencies", _vp * NSTATE),
    ]


class Adjoint(ctypes.Structure):
    _fields_ = [
        ("tendencies", _vp * NSTATE),
        ("state", _vp * NSTATE),
        ("temperature", _vp),
        ("pressure", _vp),
        ("vapor_mixing_ratio", _vp),
    ]


class WaterOutputs(ctypes.Structure):
    _fields_ = [
        ("dgncur_awet", _vp * 4),
        ("qaerwat", _vp * 4),
        ("wetdens", _vp * 4),
    ]


class WetSizeOutputs(ctypes.Structure):
    _fields_ = [
        ("dgncur_i", _vp * 4),
        ("dgncur_c", _vp * 4),
        ("dgncur_awet", _vp * 4),
        ("qaerwat", _vp * 4),
        ("wetdens", _vp * 4),
        ("tendencies", _vp * NSTATE),
    ]


class Outputs(ctypes.Structure):
    _fields_ = [
        ("tendencies", _vp * NSTATE),
        ("dgncur_i", _vp * NMODE),
        ("dgncur_c", _vp * NMODE),
        ("state_out", _vp * NSTATE),
        ("process_tendencies", _vp * NPROCESS),
        ("process_stride", ctypes.c_longlong),
        ("soa_capped", _vp),
        ("column_budgets", _vp),
    ]


class Tracers(ctypes.Structure):
    _fields_ = [
        ("state_q", _vp),
        ("qqcw", _vp),
        ("col_stride", ctypes.c_longlong),
        ("lev_stride", ctypes.c_longlong),
    ]


class Fields(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("field", _vp * NSTATE),
    ]


class SetsoxArgs(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("pressure", _vp),
        ("temperature", _vp),
        ("liquid_water", _vp),
        ("cloud_fraction", _vp),
        ("droplet_number", _vp),
        ("air_number_density", _vp),
        ("qcw", _vp * GAS_PCNST),
        ("qin", _vp * GAS_PCNST),
        ("dqdt_aqso4", _vp * GAS_PCNST),
        ("dqdt_aqh2so4", _vp * GAS_PCNST),
    ]


EXTCNT = 9
NFS = 8


class GasChemArgs(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("temperature", _vp),
        ("photolysis_rate", _vp),
        ("extfrc", _vp * EXTCNT),
        ("invariants", _vp * NFS),
        ("het_rates", _vp * GAS_PCNST),
        ("qq", _vp * GAS_PCNST),
    ]


NLINOZ = 8


class SetinvArgs(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("temperature", _vp),
        ("pressure", _vp),
        ("vapor_mixing_ratio", _vp),
        ("cnst_offline", _vp * 4),
        ("invariants", _vp * NFS),
    ]


class LinozArgs(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("temperature", _vp),
        ("pressure", _vp),
        ("pdel", _vp),
        ("climatology", _vp * NLINOZ),
        ("zenith_angle", _vp),
        ("latitude", _vp),
        ("ltrop", _vp),
        ("column_above_top", _vp),
        ("chlorine_loading", ctypes.c_double),
        ("psc_T", ctypes.c_double),
        ("o3_sfc", ctypes.c_double),
        ("o3_tau", ctypes.c_double),
        ("o3_lbl", ctypes.c_int),
        ("o3", _vp),
        ("o3_column", _vp),
        ("do3_linoz", _vp),
        ("do3_linoz_psc", _vp),
        ("ss_o3", _vp),
        ("o3_sfcsink", _vp),
    ]


PHOTO_TABLE_DIMS = ("nw", "nt", "np_xs", "numj", "nump", "numsza", "numcolo3", "numalb")
PHOTO_TABLE_ARRAYS = ("xsqy", "sza", "del_sza", "alb", "del_alb", "press", "del_p", "colo3",
                      "o3rat", "del_o3rat", "etfphot", "prs", "dprs", "rsf_tab")


class PhotoTable(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in PHOTO_TABLE_DIMS] + \
        [(n, _vp) for n in PHOTO_TABLE_ARRAYS] + \
        [("pht_alias_mult", ctypes.c_double), ("lng_index", ctypes.c_int)]


class PhotoArgs(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("pressure", _vp),
        ("pdel", _vp),
        ("temperature", _vp),
        ("o3_column", _vp),
        ("liquid_water", _vp),
        ("cloud_fraction", _vp),
        ("zenith_angle", _vp),
        ("surface_albedo", _vp),
        ("esfact", ctypes.c_double),
        ("photo", _vp),
        ("work", _vp),
    ]


class SethetArgs(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("pressure", _vp),
        ("zmid", _vp),
        ("temperature", _vp),
        ("cmfdqr", _vp),
        ("nrain", _vp),
        ("nevapr", _vp),
        ("air_density", _vp),
        ("h2o2", _vp),
        ("latitude", _vp),
        ("phis", _vp),
        ("het_rates", _vp * GAS_PCNST),
    ]


MAX_SECTORS = 4


class SetextArgs(ctypes.Structure):
    _fields_ = [
        ("ncol", ctypes.c_int),
        ("nlev", ctypes.c_int),
        ("ldc", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("frc_ndx", ctypes.c_int * EXTCNT),
        ("nsectors", ctypes.c_int * EXTCNT),
        ("file_alt_data", ctypes.c_int * EXTCNT),
        ("sectors", (_vp * MAX_SECTORS) * EXTCNT),
        ("extfrc", _vp * EXTCNT),
    ]


# entry points declared in include/mam4x.h, with ctypes signatures
SIGNATURES = {
    "mam4x_abi_version": (ctypes.c_int, []),
    "mam4x_default_config": (ctypes.c_int, [ctypes.POINTER(Config)]),
    "mam4x_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
    "mam4x_destroy": (ctypes.c_int, [_vp]),
    "mam4x_get_config": (ctypes.c_int, [_vp, ctypes.POINTER(Config)]),
    "mam4x_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mam4x_init": (ctypes.c_int, [_vp, ctypes.POINTER(Config)]),
    "mam4x_last_error": (ctypes.c_char_p, [_vp]),
    "mam4x_validate": (ctypes.c_int, [_vp, ctypes.POINTER(Inputs),
                                      ctypes.POINTER(ctypes.c_longlong), _vp]),
    "mam4x_compute_tendencies": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                                ctypes.POINTER(Inputs), ctypes.POINTER(Outputs),
                                                _vp]),
    "mam4x_compute_tendencies_noalloc": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                                        ctypes.POINTER(Inputs),
                                                        ctypes.POINTER(Outputs), _vp]),
    "mam4x_compute_tangent": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                             ctypes.POINTER(Inputs), ctypes.POINTER(Tangent),
                                             ctypes.POINTER(Outputs), _vp]),
    "mam4x_compute_adjoint": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                             ctypes.POINTER(Inputs), ctypes.POINTER(Adjoint),
                                             _vp]),
    "mam4x_compute_water_uptake": (ctypes.c_int, [_vp, ctypes.POINTER(Inputs),
                                                  ctypes.POINTER(WaterOutputs), _vp]),
    "mam4x_compute_wet_size": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.POINTER(Inputs),
                                              ctypes.POINTER(WetSizeOutputs), _vp]),
    "mam4x_tracers_to_fields": (ctypes.c_int, [_vp, ctypes.POINTER(Tracers),
                                               ctypes.POINTER(Fields), _vp]),
    "mam4x_fields_to_tracers": (ctypes.c_int, [_vp, ctypes.POINTER(Fields),
                                               ctypes.POINTER(Tracers), ctypes.c_int, _vp]),
    "mam4x_compute_gas_chemistry": (ctypes.c_int, [_vp, ctypes.c_double,
                                                   ctypes.POINTER(GasChemArgs), _vp]),
    "mam4x_compute_invariants": (ctypes.c_int, [_vp, ctypes.POINTER(SetinvArgs), _vp]),
    "mam4x_compute_linoz": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.POINTER(LinozArgs), _vp]),
    "mam4x_compute_photolysis": (ctypes.c_int, [_vp, ctypes.POINTER(PhotoTable),
                                                ctypes.POINTER(PhotoArgs), _vp]),
    "mam4x_photo_work_len": (ctypes.c_longlong, [ctypes.POINTER(PhotoTable), ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "mam4x_compute_washout": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.POINTER(SethetArgs),
                                             _vp]),
    "mam4x_compute_forcing": (ctypes.c_int, [_vp, ctypes.POINTER(SetextArgs), _vp]),
    "mam4x_compute_setsox": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.POINTER(SetsoxArgs),
                                            _vp]),
    "mam4x_batch_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(Inputs),
                                          ctypes.POINTER(Outputs), ctypes.c_double,
                                          ctypes.POINTER(_vp)]),
    "mam4x_batch_run": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp]),
    "mam4x_batch_destroy": (ctypes.c_int, [_vp]),
    "mam4x_enable_stats": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mam4x_last_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong), _vp]),
    "mam4x_last_phase_times": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_longlong), _vp]),
}

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmam4x.so")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load(path=None):
    """Load libmam4x.so (never falls back to anything else).

    torch's GPU runtime is initialised first: a process that starts a HIP
    runtime of its own before torch's can leave torch without a device."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    import torch
    torch.cuda.is_available()
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise LibraryMissing(
            "%s not built: run __graft_entry__.build() (hipcc, gfx950) first" % p)
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
