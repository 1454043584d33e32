"""NVT with the reference call surface (nep_simulator/NVT.py:9-129).

integrate_step_two / integrate_step_one and their parts call the FP64
integrator kernel part by part (nepx_nvt_part); Sim.step's fast path runs the
same arithmetic as one fused kernel (nepx_nvt_step).
"""
from . import _lib
from .Sim import Sim


class NVT(Sim):
    def integrate_step_one(self):
        self.translate1()
        self.rotate1()
        self.advance_thermostat1()

    def integrate_step_two(self):
        """Translation and torque half kicks with the thermostat factors of this step (NVT.py:17-64)."""
        self.arena.nvt_part(_lib.NVT_STEP_TWO)
        self.timestep += 1

    def translate1(self):
        self.arena.nvt_part(_lib.NVT_TRANSLATE)

    def rotate1(self):
        self.arena.nvt_part(_lib.NVT_ROTATE)

    def advance_thermostat1(self):
        """Kinetic energies, then the thermostat update the next step uses (NVT.py:108-129)."""
        self.arena.nvt_part(_lib.NVT_THERMO)
        th = self.arena.thermo()[0] if self.is_sync else None
        if th is not None:
            self.trans_kin_en = float(th["Kt"])
            self.rot_kin_en = float(th["Kr"])

