"""Translate classes of the hot-path components: the serialized field names,
start indices and output extents of the reference's translate tests
(tests/savepoint/translate/translate_fvtp2d.py, translate_riem_solver3.py,
translate_riem_solver_c.py, translate_pk3_halo.py), each running this
build's GPU component."""
from __future__ import annotations

from typing import Dict, Type

from .._config import RiemannConfig
from ..stencils.fvtp2d import FiniteVolumeTransport, PreAllocatedCopiedCornersFactory
from ..stencils.pk3_halo import PK3Halo
from ..stencils.riem_solver3 import RiemannSolver3
from ..stencils.riem_solver_c import RiemannSolverC
from ..utils import gt4py_utils as utils
from ..utils.stencil import X_DIM, Y_DIM, Z_DIM
from .translate import TranslateFortranData2Py, TranslateGrid


class TranslateFvTp2d(TranslateFortranData2Py):
    """FiniteVolumeTransport with mass fluxes and damping (translate_fvtp2d.py:10-68)."""

    def __init__(self, grid: TranslateGrid, namelist=None):
        super().__init__(grid)
        self.in_vars["data_vars"] = {
            "q": {},
            "mass": {},
            "damp_c": {},
            "nord": {"serialname": "nord_column"},
            "crx": {"istart": grid.is_},
            "cry": {"jstart": grid.js},
            "x_area_flux": {"istart": grid.is_, "serialname": "xfx"},
            "y_area_flux": {"jstart": grid.js, "serialname": "yfx"},
            "x_mass_flux": {**grid.x3d_compute_dict(), "serialname": "mfx"},
            "y_mass_flux": {**grid.y3d_compute_dict(), "serialname": "mfy"},
        }
        self.in_vars["parameters"] = ["hord"]
        self.out_vars = {
            "q": {},
            "q_x_flux": {**grid.x3d_compute_dict(), "serialname": "fx"},
            "q_y_flux": {**grid.y3d_compute_dict(), "serialname": "fy"},
        }

    def compute_from_storage(self, inputs):
        inputs["q_x_flux"] = utils.zeros(self.maxshape)
        inputs["q_y_flux"] = utils.zeros(self.maxshape)
        inputs.setdefault("mass", None)
        nord = inputs.pop("nord", None)
        damp_c = inputs.pop("damp_c", None)
        fvt = FiniteVolumeTransport(
            stencil_factory=self.grid.stencil_factory, grid_data=self.grid.grid_data,
            damping_coefficients=self.grid.damping_coefficients, grid_type=self.grid.grid_type,
            hord=int(inputs.pop("hord")), nord=nord, damp_c=damp_c)
        q_storage = inputs["q"]
        factory = PreAllocatedCopiedCornersFactory(
            self.grid.stencil_factory, dims=[X_DIM, Y_DIM, Z_DIM], y_temporary=None)
        inputs["q"] = factory(q_storage)
        fvt(**inputs)
        inputs["q"] = q_storage
        return inputs


class TranslateFvTp2d_2(TranslateFvTp2d):
    """FvTp2d without mass or mass fluxes (translate_fvtp2d.py:71-76)."""

    def __init__(self, grid: TranslateGrid, namelist=None):
        super().__init__(grid, namelist)
        for name in ("mass", "x_mass_flux", "y_mass_flux"):
            del self.in_vars["data_vars"][name]


class TranslateRiem_Solver3(TranslateFortranData2Py):
    """RiemannSolver3 (translate_riem_solver3.py)."""

    def __init__(self, grid: TranslateGrid, namelist):
        super().__init__(grid)
        self.compute_func = RiemannSolver3(
            grid.stencil_factory,
            RiemannConfig(p_fac=namelist.p_fac, a_imp=namelist.a_imp, use_logp=namelist.use_logp,
                          beta=namelist.beta))
        self.in_vars["data_vars"] = {
            "cappa": {}, "zs": {}, "w": {}, "delz": {}, "q_con": {}, "delp": {}, "pt": {},
            "zh": {},
            "pe": {"istart": grid.is_ - 1, "jstart": grid.js - 1, "kaxis": 1},
            "ppe": {}, "pk3": {}, "pk": {},
            "peln": {"istart": grid.is_, "jstart": grid.js, "kaxis": 1},
            "wsd": {"istart": grid.is_, "jstart": grid.js},
        }
        self.in_vars["parameters"] = ["dt", "ptop", "last_call"]
        self.out_vars = {
            "zh": {"kend": grid.npz},
            "w": {},
            "pe": {"istart": grid.is_ - 1, "iend": grid.ie + 1, "jstart": grid.js - 1,
                   "jend": grid.je + 1, "kend": grid.npz, "kaxis": 1},
            "peln": {"istart": grid.is_, "iend": grid.ie, "jstart": grid.js, "jend": grid.je,
                     "kend": grid.npz, "kaxis": 1},
            "ppe": {"kend": grid.npz},
            "delz": {},
            "pk": grid.compute_buffer_k_dict(),
            "pk3": grid.default_buffer_k_dict(),
        }

    def compute_from_storage(self, inputs):
        inputs["last_call"] = bool(inputs["last_call"])
        self.compute_func(**inputs)
        return inputs


class TranslateRiem_Solver_C(TranslateFortranData2Py):
    """RiemannSolverC (translate_riem_solver_c.py)."""

    def __init__(self, grid: TranslateGrid, namelist):
        super().__init__(grid)
        self.compute_func = RiemannSolverC(grid.stencil_factory, namelist.p_fac)
        self.in_vars["data_vars"] = {"cappa": {}, "hs": {}, "w3": {}, "ptc": {}, "q_con": {},
                                     "delpc": {}, "gz": {}, "pef": {}, "ws": {}}
        self.in_vars["parameters"] = ["dt2", "ptop"]
        self.out_vars = {"pef": {"kend": grid.npz}, "gz": {"kend": grid.npz}}
        self.max_error = 5e-14


class TranslatePK3_Halo(TranslateFortranData2Py):
    """PK3Halo (translate_pk3_halo.py)."""

    def __init__(self, grid: TranslateGrid, namelist=None):
        super().__init__(grid)
        self.compute_func = PK3Halo(grid.stencil_factory)
        self.in_vars["data_vars"] = {"pk3": {}, "delp": {}}
        self.in_vars["parameters"] = ["akap", "ptop"]
        self.out_vars = {"pk3": {"kend": grid.npz + 1}}


# savepoint test name -> translate class
TRANSLATE_CLASSES: Dict[str, Type[TranslateFortranData2Py]] = {
    "FvTp2d": TranslateFvTp2d,
    "FvTp2d_2": TranslateFvTp2d_2,
    "Riem_Solver3": TranslateRiem_Solver3,
    "Riem_Solver_C": TranslateRiem_Solver_C,
    "PK3_Halo": TranslatePK3_Halo,
}
