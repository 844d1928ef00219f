"""ctypes mirror of the C-ABI types declared in include/nemo_hip.h.

The same structures are handed to the HIP library and to the CPU oracle, so
their field order must match the header exactly; tests check the sizes.
"""
import ctypes as C

NEMO_HIP_OK = 0
NEMO_HIP_ERR_ARG = 1
NEMO_HIP_ERR_UNSUPPORTED = 2
NEMO_HIP_ERR_STATE = 3
NEMO_HIP_ERR_HIP = 4
NEMO_HIP_ERR_NUMERIC = 5

NEMO_HIP_PTR_HOST = 0
NEMO_HIP_PTR_DEVICE = 1

# vorticity schemes, DYN/dynvor.F90:65-70
NP_ENS, NP_ENE, NP_ENT, NP_EET, NP_EEN, NP_MIX = range(6)
VOR_NAMES = {"ens": NP_ENS, "ene": NP_ENE, "ent": NP_ENT, "eet": NP_EET, "een": NP_EEN, "mix": NP_MIX}
# momentum advection forms, DYN/dynadv.F90:41-44
NP_LIN_DYN, NP_VEC_C2, NP_FLX_C2, NP_FLX_UBS = range(4)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

DOMAIN_INT_FIELDS = [
    "jpi", "jpj", "jpk", "jperio",
    "ln_linssh", "ln_isfcav", "ln_zps", "ln_sco",
    "ln_traldf_lap", "ln_traldf_msc", "ln_zdfddm", "ln_zad_Aimp",
    "n_dynadv", "nn_dynkeg", "nvor_scheme", "nn_een_e3f",
    "ln_drgimp",
    "ln_dynspg_ts", "ln_bt_fw", "ln_bt_av", "nn_bt_flt", "nn_baro",
    "neuler", "nit000", "ln_traqsr",
]
DOMAIN_REAL_FIELDS = ["rdt", "rn_atfp", "rau0", "grav", "rn_bt_alpha", "rn_scal_load"]
DOMAIN_KIDX_FIELDS = ["k_top", "k_bot"]
DOMAIN_2D_FIELDS = [
    "e1t", "e2t", "e1u", "e2u", "e1v", "e2v", "e1f", "e2f",
    "e1e2t", "r1_e1e2t", "e1e2u", "r1_e1e2u", "e1e2v", "r1_e1e2v",
    "e1e2f", "r1_e1e2f",
    "e2_e1u", "e1_e2v", "r1_e1u", "r1_e2u", "r1_e1v", "r1_e2v",
    "ff_f", "ff_t",
]
DOMAIN_3D_FIELDS = ["e3t_n", "e3u_n", "e3v_n", "e3w_n", "e3f_n", "e3uw_n", "e3vw_n"]


class DomainStruct(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in DOMAIN_INT_FIELDS]
                + [(n, C.c_double) for n in DOMAIN_REAL_FIELDS]
                + [(n, _ip) for n in DOMAIN_KIDX_FIELDS]
                + [(n, _dp) for n in DOMAIN_2D_FIELDS]
                + [(n, _dp) for n in DOMAIN_3D_FIELDS])


SPG_3D_FIELDS = ["ua", "va", "un", "vn"]
SPG_IN_FIELDS = ["un_b", "vn_b", "ub_b", "vb_b", "sshn", "hu_n", "hv_n", "r1_hu_n", "r1_hv_n", "ht_n",
                 "rCdU_bot", "utau", "vtau", "emp", "rnf"]
SPG_OUT_FIELDS = ["ua_b", "va_b", "ssha", "un_adv", "vn_adv"]
SPG_STATE_FIELDS = ["ub2_b", "vb2_b", "un_bf", "vn_bf"]


class SpgTsStruct(C.Structure):
    _fields_ = ([(n, C.c_void_p) for n in SPG_3D_FIELDS]
                + [(n, C.c_void_p) for n in SPG_IN_FIELDS + SPG_OUT_FIELDS + SPG_STATE_FIELDS])


# equation of state, TRA/eosbn2.F90:84-95 (nameos defaults of the S-EOS)
NP_TEOS10, NP_EOS80, NP_SEOS = -1, 0, 1
EOS_REAL_FIELDS = ["rn_a0", "rn_b0", "rn_lambda1", "rn_lambda2", "rn_mu1", "rn_mu2", "rn_nu"]
SEOS_DEFAULTS = dict(rn_a0=1.6550e-1, rn_b0=7.6554e-1, rn_lambda1=5.9520e-2, rn_lambda2=5.4914e-4,
                     rn_mu1=1.4970e-4, rn_mu2=1.1090e-5, rn_nu=2.4341e-3)


class EosStruct(C.Structure):
    _fields_ = [("neos", C.c_int32), ("reserved0", C.c_int32)] + [(n, C.c_double) for n in EOS_REAL_FIELDS]


def seos(**kw):
    """An EosStruct for the simplified EOS (defaults of TRA/eosbn2.F90:89-95)."""
    e = EosStruct()
    e.neos = NP_SEOS
    v = dict(SEOS_DEFAULTS)
    v.update(kw)
    for n in EOS_REAL_FIELDS:
        setattr(e, n, float(v[n]))
    return e


def teos10():
    """An EosStruct selecting the polynomial TEOS-10 (ln_TEOS10, TRA/eosbn2.F90:1266)."""
    e = seos()
    e.neos = NP_TEOS10
    return e


def eos80():
    """An EosStruct selecting the polynomial EOS-80 (ln_EOS80, TRA/eosbn2.F90:1457)."""
    e = seos()
    e.neos = NP_EOS80
    return e


# restart state, nemo_hip_rst_t (IOM/restart.F90:133-309, DYN/dynspg_ts.F90:1391-1463)
RST_FIELDS = ["ub", "vb", "tsb", "un", "vn", "tsn", "rhop", "sshb", "sshn", "ub2_b", "vb2_b", "un_bf", "vn_bf",
              "un_b", "vn_b", "ub_b", "vb_b", "r1_hu_n", "r1_hv_n", "qsr_hc", "qsr_hc_b", "fraqsr_1lev"]


class RstStruct(C.Structure):
    _fields_ = [("real_bytes", C.c_int32), ("kqsr_b", C.c_int32)] + [(n, C.c_void_p) for n in RST_FIELDS]


# dia_hsb state and results (nemo_hip_hsb_t)
HSB_PTR_FIELDS = ["tsn", "tsb", "wn", "sshn", "emp", "rnf", "qsr", "sbc_tsc", "surf_ini", "ssh_ini",
                  "ssh_hc_loc_ini", "ssh_sc_loc_ini", "e3t_ini", "hc_loc_ini", "sc_loc_ini"]
HSB_SCALARS = ["surf_tot", "frc_v", "frc_t", "frc_s", "frc_wn_t", "frc_wn_s"]
HSB_DIAG = ["bgfrcvol", "bgfrctem", "bgfrchfx", "bgfrcsal", "bgtemper", "bgsaline", "bgheatco", "bgheatfx",
            "bgsaltco", "bgvolssh", "bgmistem", "bgmissal"]


class HsbStruct(C.Structure):
    _fields_ = ([("real_bytes", C.c_int32), ("reserved0", C.c_int32)]
                + [(n, C.c_void_p) for n in HSB_PTR_FIELDS]
                + [(n, C.c_double) for n in HSB_SCALARS]
                + [("sums", C.c_double * 13)]
                + [(n, C.c_double) for n in HSB_DIAG])
