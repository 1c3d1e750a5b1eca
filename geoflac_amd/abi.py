"""ctypes mirror of include/gf_abi.h (structs, field bits, status codes).

The structures follow the header field for field; `check_layout()` compares
their sizes with the values the compiled library reports, so a drift between
the header and this file fails loudly instead of corrupting memory.
"""
import ctypes as C

MAXPH = 20      # src/params.f90:9
MAXINH = 100    # src/params.f90:12

GF_OK = 0
GF_STOP_MAXWELL = 22
GF_STOP_NEGATIVE_AREA = 40
GF_ERR_HIP = 1000
GF_ERR_ARG = 1001
GF_ERR_UNSUPPORTED = 1002
GF_ERR_STATE = 1003
GF_ERR_THERM_SWEEPS = 1004
GF_ERR_PHASES = 1005

PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_STORAGE = 2

_I = C.c_int32
_D = C.c_double

_INT_FIELDS = [
    "nx", "nz", "nphase",
    "mix_strain", "mix_stress", "ndim", "movegrid",
    "itherm", "istress_therm", "ishearh", "itemp_bc", "itype_melting",
    "nystressbc", "nydrsides", "nyhydro", "i_prestress",
    "idt_scale", "i_rey", "ivis_present", "iphsub",
    "ifreq_rmasses", "ifreq_imasses", "ifreq_visc", "ifreq_avgsr",
    "inhom",
]
_REAL_FIELDS = [
    "t_top", "t_bot", "bot_bc", "g", "demf", "ten_off", "tau_heal", "v_min", "v_max",
    "frac", "fracm", "strain_inert", "dt_scale", "xReyn", "vbc", "rxbo", "rzbo",
    "pisos", "drosub", "fmagma_max", "weaken_ratio_plastic", "weaken_ratio_viscous", "rho_magma",
]
PHASE_TABLES = [
    "den", "alfa", "beta", "pln", "acoef", "eactiv", "vactiv", "rl", "rm",
    "plstrain1", "plstrain2", "fric1", "fric2", "cohesion1", "cohesion2",
    "dilat1", "dilat2", "conduct", "cp",
]


class GfParams(C.Structure):
    _fields_ = ([(n, _I) for n in _INT_FIELDS]
                + [("irheol", _I * MAXPH)]
                + [(n, _I * MAXINH) for n in ("ix1", "ix2", "iy1", "iy2", "igeom")]
                + [("flexure_base", _I), ("nelem_serp", _I), ("nopbmax", _I)]
                + [(n, _D) for n in _REAL_FIELDS]
                + [(n, _D * MAXPH) for n in PHASE_TABLES]
                + [("topo_kappa", _D), ("erodibility", _D * MAXPH)])


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_FP = C.POINTER(C.c_float)
_BP = C.POINTER(C.c_int8)

# surface processes (topo_kappa > 0): dtopo(nx), dhacc(nx-1), dhacc_correction(nx)
TOPO_FIELDS = [("dtopo", _DP), ("dhacc", _DP), ("dhacc_correction", _DP)]

HOST_FIELDS = [
    ("cord", _DP), ("vel", _DP), ("temp", _DP), ("temp0", _DP), ("force", _DP),
    ("amass", _DP), ("rmass", _DP), ("bc", _DP), ("ncod", _IP),
    ("stress0", _DP), ("area", _DP), ("dvol", _DP), ("strain", _DP), ("se2sr", _DP),
    ("aps", _DP), ("visn", _DP), ("e2sr", _DP), ("sshrheat", _DP), ("shrheat", _DP),
    ("source", _DP), ("iphase", _IP), ("phase_ratio", _DP),
]


class GfHostArrays(C.Structure):
    _fields_ = HOST_FIELDS + [("nopbou", _IP), ("bcstress", _DP), ("coh_scale", _FP)] + TOPO_FIELDS


MAX_MARKERS_PER_ELEM = 32  # src/marker_data.f90:5
MARKER_FIELDS = [("mark_a1", _DP), ("mark_a2", _DP), ("mark_ntriag", _IP), ("mark_phase", _BP),
                 ("mark_dead", _BP), ("mark_id_elem", _IP), ("nmark_elem", _IP)]


class GfMarkerArrays(C.Structure):
    _fields_ = [("nmarkers", _I), ("capacity", _I)] + MARKER_FIELDS + [("mark_age", _DP)]


class GfClock(C.Structure):
    _fields_ = [("nloop", _I), ("nsrate", _I), ("time", _D), ("dt", _D), ("dtavg", _D),
                ("dt_elastic", _D), ("dt_maxwell", _D)]


IO_FLAGS = ["io_vel", "io_srII", "io_eII", "io_aps", "io_sII", "io_sxx", "io_szz", "io_sxz", "io_pres", "io_temp",
            "io_phase", "io_visc", "io_unused", "io_density", "io_src", "io_diss", "io_forc", "io_hfl", "io_topo"]
# host-owned arrays of the output files (none: zeros)
OUTPUT_ARRAYS = ["xoriginal", "zoriginal", "dhacc", "extr_acc", "dhacc_correction", "q_init", "vel_flex_old",
                 "fmagma", "fmagma2", "dike_backlog", "dike_marker_vol", "mor_marker_vol", "mark_age"]


class GfOutputOpts(C.Structure):
    _fields_ = ([(n, _I) for n in IO_FLAGS] + [("lastout", _I), ("hs", _D)]
                + [(n, _DP) for n in OUTPUT_ARRAYS])


# field bits, in the order of HOST_FIELDS
FIELD_BIT = {name: 1 << k for k, (name, _) in enumerate(HOST_FIELDS)}
GF_ALL = (1 << len(HOST_FIELDS)) - 1
GF_MARKERS = 1 << 22
GF_COH_SCALE = 1 << 23
GF_STRESSBC = 1 << 24
GF_TOPO = 1 << 25

SIZEOF_PARAMS = C.sizeof(GfParams)
SIZEOF_HOST = C.sizeof(GfHostArrays)
SIZEOF_CLOCK = C.sizeof(GfClock)
