"""Deck values and the derived nondimensional constants.

OrgDeck holds the value lines of original.org in the order INPUT_ReadOrg reads them
(Input/ReadOrg.F90:106-151); GlobalInfo holds the turbomachinery lines of
dimension_global.dat (Input/Read_GlobalInfo.F90:209-214).  make_params() performs
ParameterSet and NonDimensionlization (Input/ParameterSet.F90:1-96) and returns the
trans_params struct of the C-ABI.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, fields

from . import capi


@dataclass
class OrgDeck:
    KR: int = 0
    MA: int = 2000
    Kind_RotMod: int = 0
    IF_parallel: int = 0
    KO: int = 1
    KQ: int = 100000
    Num_dat: int = 0
    kind_field: int = 0
    if_VTK: int = 0
    IF_Dst: int = 0
    Alscale: float = 1.0
    XWc: float = 0.0
    YWc: float = 0.0
    ZWc: float = 0.0
    Area_Ref: float = 1.0
    C_Ref: float = 1.0
    B_Ref: float = 1.0
    XM: float = 0.5
    AoA: float = 0.0
    AoYaw: float = 0.0
    RefOrg: float = 0.0
    Tinf: float = 288.0
    If_EquT: int = 0
    Wall_Temp: float = 288.0
    Delta_BL: float = 0.0
    IF_NS_Uniform: int = 1
    Q_Dst: float = 0.0
    VisRatio: float = 1.0
    FSTI: float = 0.01
    Kind_Spatial: int = 2
    Kind_Roe: int = 1
    E2: float = 1.0
    E4: float = 1.0
    F_limiter1: float = 0.0
    F_limiter2: float = 0.0
    gama_max: float = 0.5
    gama_min: float = 0.01
    IF_Precondition: int = 0
    DT_LUSGS_Main: float = 1.0e10
    CFL_Sub_start: float = 5.0
    CFL_Sub_end: float = 5.0
    CFL_Sub_step: int = 1
    Kind_Dual: int = 2
    Kchld: int = 1
    IF_RK: int = 0
    IF_Euler: int = 1
    IF_Turb_Org: int = 0
    IF_setRkof: int = 0
    Rkfset: float = 0.0
    Rofset: float = 0.0
    KItm: int = 1
    Kind_Model: int = 1
    Prod_Lmt: float = 20.0
    If_wallF: int = 0
    IF_Comppress: int = 0
    Kind_CC: int = 0
    Kind_Hybrid: int = 0
    Kind_SubHY: int = 0
    IF_TRANSITION_org: int = 0
    Tuinf: float = 0.0
    RmiuLmax: float = 0.0
    If_fix_Transition: int = 0
    Trans_location: float = 0.0
    Krr: int = 0
    Ki_Aver: int = 10 ** 8
    KI_statis: int = 10 ** 8
    KI_C: int = 10 ** 8
    KSss: int = 10 ** 8


# value lines of original.org, in file order (Input/ReadOrg.F90:109-151)
ORG_LINES = [
    ["KR", "MA", "Kind_RotMod"],
    ["IF_parallel", "KO", "KQ", "Num_dat", "kind_field", "if_VTK"],
    ["IF_Dst", "Alscale", "XWc", "YWc", "ZWc", "Area_Ref", "C_Ref", "B_Ref"],
    ["XM", "AoA", "AoYaw", "RefOrg", "Tinf", "If_EquT", "Wall_Temp", "Delta_BL"],
    ["IF_NS_Uniform", "Q_Dst", "VisRatio", "FSTI"],
    ["Kind_Spatial", "Kind_Roe", "E2", "E4", "F_limiter1", "F_limiter2", "gama_max",
     "gama_min", "IF_Precondition"],
    ["DT_LUSGS_Main", "CFL_Sub_start", "CFL_Sub_end", "CFL_Sub_step"],
    ["Kind_Dual", "Kchld", "IF_RK"],
    ["IF_Euler", "IF_Turb_Org", "IF_setRkof", "Rkfset", "Rofset"],
    ["KItm", "Kind_Model", "Prod_Lmt", "If_wallF"],
    ["IF_Comppress", "Kind_CC"],
    ["Kind_Hybrid", "Kind_SubHY"],
    ["IF_TRANSITION_org", "Tuinf", "RmiuLmax"],
    ["If_fix_Transition", "Trans_location"],
    ["Krr", "Ki_Aver", "KI_statis", "KI_C", "KSss"],
]


@dataclass
class GlobalInfo:
    nblocks: int = 1
    P0in: float = 101325.0
    T0in: float = 288.0
    Pouthub: float = 101325.0
    nblade: int = 1
    Npassage: int = 1
    omega_rpm: tuple = (0.0, 0.0, 0.0)
    Rot_Ori: tuple = (0.0, 0.0, 0.0)


def make_params(org: OrgDeck, glob: GlobalInfo, implicit_fp32: int = 0) -> capi.TransParams:
    p = capi.TransParams()
    pi = 4.0 * math.atan(1.0)
    alfa = org.AoA / 180.0 * pi
    beta = org.AoYaw / 180.0 * pi
    xm = org.XM
    gam0 = 1.4
    xm2 = xm * xm
    rxm2 = xm2 * gam0
    ppf = 1.0 / gam0 / xm2
    csthlnd = 110.0 / org.Tinf
    csth1 = org.Tinf / 273.15
    # turbulence constants (ParameterSet.F90:32-65)
    a1, akapa, beta_star = 0.31, 0.41, 0.09
    sigk1, sigo1, beta1, gama1 = 0.85, 0.5, 0.075, 5.0 / 9.0
    sigk2, sigo2, beta2, gama2 = 1.0, 0.856, 0.0828, 0.44
    rmiuf = 1.8e-5 / 1.2
    rkf = 1.5 * org.FSTI ** 2.0
    rof = 1.2 * rkf / rmiuf
    if org.IF_setRkof == 1:
        rkf, rof = org.Rkfset, org.Rofset
    # NonDimensionlization (ParameterSet.F90:70-96)
    t0in = glob.T0in / org.Tinf
    asonic = math.sqrt(1.4 * org.Tinf * 287)
    vref = xm * asonic
    lref = 1.0
    miuref = (1.0 + csthlnd) / (t0in * csth1 + csthlnd) * (csth1 * t0in) ** 1.5 * 1.7e-5
    roref = 1.2
    ref = roref * vref * lref / miuref
    pref = roref * vref ** 2.0
    p0in = glob.P0in / pref
    pouthub = glob.Pouthub / pref
    omega1 = glob.omega_rpm[0] * 2.0 * pi * lref / (60.0 * vref)
    comm = 1.0 + 0.2 * xm * xm
    ps00 = p0in / comm ** 3.5
    ts00 = t0in / comm

    p.kind_spatial, p.kind_roe, p.if_euler = org.Kind_Spatial, org.Kind_Roe, org.IF_Euler
    p.if_turb_org, p.kind_model = org.IF_Turb_Org, org.Kind_Model
    p.kind_dual, p.kchld, p.if_rk, p.kitm = org.Kind_Dual, org.Kchld, org.IF_RK, org.KItm
    p.kind_hybrid, p.kind_subhy = org.Kind_Hybrid, org.Kind_SubHY
    p.if_comppress, p.kind_cc = org.IF_Comppress, org.Kind_CC
    p.kind_rotmod, p.if_precondition = org.Kind_RotMod, org.IF_Precondition
    p.cfl_sub_step, p.ki_ini, p.implicit_fp32 = org.CFL_Sub_step, 0, implicit_fp32
    p.xm, p.rxm2, p.ppf, p.csthlnd, p.csth1, p.ref = xm, rxm2, ppf, csthlnd, csth1, ref
    p.omega1, p.tiny = omega1, 1.0e-20
    p.vxf = math.cos(beta) * math.cos(alfa)
    p.vyf = math.cos(beta) * math.sin(alfa)
    p.vzf = math.sin(beta)
    p.dt_lusgs_main, p.cfl_sub_start, p.cfl_sub_end = (org.DT_LUSGS_Main, org.CFL_Sub_start,
                                                       org.CFL_Sub_end)
    p.gama_max, p.gama_min = org.gama_max, org.gama_min
    p.prod_lmt, p.visratio, p.fsti = org.Prod_Lmt, org.VisRatio, org.FSTI
    p.a1, p.akapa, p.beta_star = a1, akapa, beta_star
    p.sigk1, p.sigo1, p.beta1, p.gama1 = sigk1, sigo1, beta1, gama1
    p.sigk2, p.sigo2, p.beta2, p.gama2 = sigk2, sigo2, beta2, gama2
    p.p0in, p.t0in, p.pouthub, p.ps00, p.ts00 = p0in, t0in, pouthub, ps00, ts00
    p.rkf, p.rof, p.rmiuf = rkf, rof, rmiuf
    p.f_limiter1, p.f_limiter2, p.alscale = org.F_limiter1, org.F_limiter2, org.Alscale
    return p


def params_dict(p: capi.TransParams) -> dict:
    return {name: getattr(p, name) for name, _ in p._fields_}


def org_fields():
    return [f.name for f in fields(OrgDeck)]
