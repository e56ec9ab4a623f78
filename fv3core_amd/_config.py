"""Namelist configuration dataclasses (fv3core/_config.py:85-431).

Field names and meaning follow the reference, so namelists written for it
configure this package unchanged.  ``DynamicalCoreConfig`` carries the
dycore namelist; the frozen per-component configs are derived from it by the
same properties (``riemann``, ``d_grid_shallow_water``, ``acoustic_dynamics``,
``remapping``).  Options that only matter to physics outside the hot path
(cloud microphysics) are not carried.
"""
from __future__ import annotations

import dataclasses
from typing import Tuple


@dataclasses.dataclass(frozen=True)
class RiemannConfig:
    p_fac: float
    a_imp: float
    use_logp: bool
    beta: float


@dataclasses.dataclass(frozen=True)
class DGridShallowWaterLagrangianDynamicsConfig:
    dddmp: float
    d2_bg: float
    d2_bg_k1: float
    d2_bg_k2: float
    d4_bg: float
    ke_bg: float
    nord: int
    n_sponge: int
    grid_type: int
    d_ext: float
    hord_dp: int
    hord_tm: int
    hord_mt: int
    hord_vt: int
    do_f3d: bool
    do_skeb: bool
    d_con: float
    vtdm4: float
    inline_q: bool
    convert_ke: bool
    do_vort_damp: bool
    hydrostatic: bool


@dataclasses.dataclass(frozen=True)
class AcousticDynamicsConfig:
    tau: float
    k_split: int
    n_split: int
    m_split: int
    delt_max: float
    rf_cutoff: float
    rf_fast: bool
    breed_vortex_inline: bool
    use_old_omega: bool
    riemann: RiemannConfig
    d_grid_shallow_water: DGridShallowWaterLagrangianDynamicsConfig

    @property
    def nord(self) -> int:
        return self.d_grid_shallow_water.nord

    @property
    def grid_type(self) -> int:
        return self.d_grid_shallow_water.grid_type

    @property
    def hydrostatic(self) -> bool:
        return self.d_grid_shallow_water.hydrostatic

    @property
    def hord_tm(self) -> int:
        return self.d_grid_shallow_water.hord_tm

    @property
    def p_fac(self) -> float:
        return self.riemann.p_fac

    @property
    def d_ext(self) -> float:
        return self.d_grid_shallow_water.d_ext

    @property
    def d_con(self) -> float:
        return self.d_grid_shallow_water.d_con

    @property
    def beta(self) -> float:
        return self.riemann.beta

    @property
    def use_logp(self) -> bool:
        return self.riemann.use_logp


@dataclasses.dataclass(frozen=True)
class SatAdjustConfig:
    """Namelist of the fast saturation adjustment (_config.py:86-110)."""

    hydrostatic: bool = False
    rad_snow: bool = True
    rad_rain: bool = True
    rad_graupel: bool = True
    tintqs: bool = False
    sat_adj0: float = 0.90
    ql_gen: float = 1.0e-3
    qs_mlt: float = 1.0e-6
    ql0_max: float = 2.0e-3
    t_sub: float = 184.0
    qi_gen: float = 1.82e-6
    qi_lim: float = 1.0
    qi0_max: float = 1.0e-4
    dw_ocean: float = 0.10
    dw_land: float = 0.20
    icloud_f: int = 0
    cld_min: float = 0.05
    tau_i2s: float = 1000.0
    tau_v2l: float = 150.0
    tau_r2g: float = 900.0
    tau_l2r: float = 900.0
    tau_l2v: float = 300.0
    tau_imlt: float = 600.0
    tau_smlt: float = 900.0


@dataclasses.dataclass(frozen=True)
class RemappingConfig:
    fill: bool
    kord_tm: int
    kord_tr: int
    kord_wz: int
    kord_mt: int
    do_sat_adj: bool
    sat_adjust: SatAdjustConfig = dataclasses.field(default_factory=SatAdjustConfig)

    @property
    def hydrostatic(self) -> bool:
        return self.sat_adjust.hydrostatic


@dataclasses.dataclass
class DynamicalCoreConfig:
    dt_atmos: int = 225
    a_imp: float = 1.0
    beta: float = 0.0
    consv_te: float = 0.0
    d2_bg: float = 0.0
    d2_bg_k1: float = 0.2
    d2_bg_k2: float = 0.1
    d4_bg: float = 0.15
    d_con: float = 1.0
    d_ext: float = 0.0
    dddmp: float = 0.5
    delt_max: float = 0.002
    do_sat_adj: bool = False
    do_vort_damp: bool = True
    fill: bool = True
    hord_dp: int = 6
    hord_mt: int = 6
    hord_tm: int = 6
    hord_tr: int = 8
    hord_vt: int = 6
    hydrostatic: bool = False
    k_split: int = 1
    ke_bg: float = 0.0
    kord_mt: int = 9
    kord_tm: int = -9
    kord_tr: int = 9
    kord_wz: int = 9
    n_split: int = 6
    nord: int = 3
    npx: int = 97
    npy: int = 97
    npz: int = 79
    ntiles: int = 6
    nwat: int = 6
    p_fac: float = 0.05
    rf_cutoff: float = 3000.0
    tau: float = 10.0
    vtdm4: float = 0.06
    z_tracer: bool = True
    do_qa: bool = True
    layout: Tuple[int, int] = (1, 1)
    grid_type: int = 0
    do_f3d: bool = False
    inline_q: bool = False
    do_skeb: bool = False
    use_logp: bool = False
    moist_phys: bool = True
    m_split: int = 0
    convert_ke: bool = False
    breed_vortex_inline: bool = False
    use_old_omega: bool = True
    rf_fast: bool = True
    p_ref: float = 1e5
    adiabatic: bool = False
    nf_omega: int = 1
    fv_sg_adj: int = 600
    n_sponge: int = 48
    check_negative: bool = False
    c2l_ord: int = 4
    # fast saturation adjustment (NamelistDefaults of the reference's _config.py)
    tau_r2g: float = 900.0
    tau_smlt: float = 900.0
    tau_imlt: float = 600.0
    tau_i2s: float = 1000.0
    tau_l2r: float = 900.0
    tau_l2v: float = 300.0
    tau_v2l: float = 150.0
    sat_adj0: float = 0.90
    ql_gen: float = 1.0e-3
    qs_mlt: float = 1.0e-6
    ql0_max: float = 2.0e-3
    t_sub: float = 184.0
    qi_gen: float = 1.82e-6
    qi_lim: float = 1.0
    qi0_max: float = 1.0e-4
    rad_snow: bool = True
    rad_rain: bool = True
    rad_graupel: bool = True
    tintqs: bool = False
    dw_ocean: float = 0.10
    dw_land: float = 0.20
    icloud_f: int = 0
    cld_min: float = 0.05

    @property
    def riemann(self) -> RiemannConfig:
        return RiemannConfig(p_fac=self.p_fac, a_imp=self.a_imp, use_logp=self.use_logp,
                             beta=self.beta)

    @property
    def d_grid_shallow_water(self) -> DGridShallowWaterLagrangianDynamicsConfig:
        return DGridShallowWaterLagrangianDynamicsConfig(
            dddmp=self.dddmp, d2_bg=self.d2_bg, d2_bg_k1=self.d2_bg_k1, d2_bg_k2=self.d2_bg_k2,
            d4_bg=self.d4_bg, ke_bg=self.ke_bg, nord=self.nord, n_sponge=self.n_sponge,
            grid_type=self.grid_type, d_ext=self.d_ext, inline_q=self.inline_q,
            hord_dp=self.hord_dp, hord_tm=self.hord_tm, hord_mt=self.hord_mt,
            hord_vt=self.hord_vt, do_f3d=self.do_f3d, do_skeb=self.do_skeb, d_con=self.d_con,
            vtdm4=self.vtdm4, do_vort_damp=self.do_vort_damp, hydrostatic=self.hydrostatic,
            convert_ke=self.convert_ke,
        )

    @property
    def acoustic_dynamics(self) -> AcousticDynamicsConfig:
        return AcousticDynamicsConfig(
            tau=self.tau, k_split=self.k_split, n_split=self.n_split, m_split=self.m_split,
            delt_max=self.delt_max, rf_fast=self.rf_fast, rf_cutoff=self.rf_cutoff,
            breed_vortex_inline=self.breed_vortex_inline, use_old_omega=self.use_old_omega,
            riemann=self.riemann, d_grid_shallow_water=self.d_grid_shallow_water,
        )

    @property
    def sat_adjust(self) -> SatAdjustConfig:
        return SatAdjustConfig(
            hydrostatic=self.hydrostatic, rad_snow=self.rad_snow, rad_rain=self.rad_rain,
            rad_graupel=self.rad_graupel, tintqs=self.tintqs, sat_adj0=self.sat_adj0,
            ql_gen=self.ql_gen, qs_mlt=self.qs_mlt, ql0_max=self.ql0_max, t_sub=self.t_sub,
            qi_gen=self.qi_gen, qi_lim=self.qi_lim, qi0_max=self.qi0_max,
            dw_ocean=self.dw_ocean, dw_land=self.dw_land, icloud_f=self.icloud_f,
            cld_min=self.cld_min, tau_i2s=self.tau_i2s, tau_v2l=self.tau_v2l,
            tau_r2g=self.tau_r2g, tau_l2r=self.tau_l2r, tau_l2v=self.tau_l2v,
            tau_imlt=self.tau_imlt, tau_smlt=self.tau_smlt)

    @property
    def remapping(self) -> RemappingConfig:
        return RemappingConfig(fill=self.fill, kord_tm=self.kord_tm, kord_tr=self.kord_tr,
                               kord_wz=self.kord_wz, kord_mt=self.kord_mt,
                               do_sat_adj=self.do_sat_adj, sat_adjust=self.sat_adjust)
