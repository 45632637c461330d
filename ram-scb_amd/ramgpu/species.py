"""RAM species table (mirrors ``DefineSpecies``, src/ModRamSpecies.f90:37-138).

Only the fields the advection step reads are kept: mass (in proton masses),
charge state, whether the species charge-exchanges, which analytic
cross-section fit ``CEPARA`` selects for it (src/ModRamLoss.f90:37-83, chosen
by ``species%s_name``), ``WPI`` (electron wave-particle lifetimes, WAVELO,
on for electrons as in the reference) and the plasmasphere composition ratio
the Coulomb parameters use.  Whether a run applies WAVELO is the problem's
choice (``inputs.make_problem(wave=True)``).
"""
from __future__ import annotations

from dataclasses import dataclass

# charge-exchange fit families, matching the C-ABI's RAMGPU_CEX_* values
CEX_NONE, CEX_H, CEX_HE, CEX_O = 0, 1, 2, 3


@dataclass(frozen=True)
class Species:
    name: str        # species%s_name
    code: str        # species%s_code
    mass: float      # species%s_mass, in units of MP
    charge: int      # species%s_charge
    cex: bool        # species%CEX
    cex_kind: int    # CEX_* fit used by CEPARA when cex is true
    wpi: bool = False
    pp_ratio: float = 0.0  # species%plasmasphereRatio


# src/ModRamSpecies.f90:43-104
ELECTRON = Species("Electron", "_e", 5.4462E-4, -1, False, CEX_NONE, wpi=True, pp_ratio=1.0)
HYDROGEN = Species("Hydrogen", "_H", 1.0, 1, True, CEX_H, pp_ratio=0.77)
HELIUM = Species("HeliumP1", "He", 4.0, 1, True, CEX_HE, pp_ratio=0.2)
OXYGEN = Species("OxygenP1", "_O", 16.0, 1, True, CEX_O, pp_ratio=0.03)

BY_CODE = {sp.code: sp for sp in (ELECTRON, HYDROGEN, HELIUM, OXYGEN)}
BY_NAME = {sp.name: sp for sp in (ELECTRON, HYDROGEN, HELIUM, OXYGEN)}
