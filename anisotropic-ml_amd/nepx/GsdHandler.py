"""GsdHandler with the reference call surface (nep_simulator/GsdHandler.py:5-131), on nepx.gsd."""
import numpy as np

from . import gsd


class GsdHandler:
    def __init__(self, filename):
        self.read_system(filename, -1)

    def read_system(self, input, target_frame):
        fr = gsd.read_frame(input, target_frame)
        if fr["_frames"] > 1:
            print("init_gsd can't have more than one frames")
        self.frame = fr["_frame"]
        self.target_frame = target_frame
        # exact FP64 state and thermostat when the file was written by nepx
        self.positions = fr.get("log/nepx/position_f64", fr["particles/position"]).copy()
        self.velocities = fr.get("log/nepx/velocity_f64", fr["particles/velocity"]).copy()
        self.orientations = fr.get("log/nepx/orientation_f64", fr["particles/orientation"]).copy()
        self.angmom = fr.get("log/nepx/angmom_f64", fr["particles/angmom"]).copy()
        self.bodi = fr["particles/body"].copy()
        self.moment_inertia = fr["particles/moment_inertia"].copy()
        self.mass = fr["particles/mass"].copy()
        self.types = fr.get("particles/types")
        self.typeid = fr["particles/typeid"].copy()
        self.charges = fr.get("particles/charge", np.zeros(len(self.typeid), dtype=np.float32)).copy()
        self.box = fr["configuration/box"]
        self.Lx, self.Ly, self.Lz = (float(v) for v in self.box[0:3])
        th = fr.get("log/nepx/thermostat")
        self.thermostat = None if th is None else dict(xi=th[0], eta=th[1], xi_r=th[2], eta_r=th[3])
        self.body = np.array(self.bodi, dtype=np.int64)
        self.body[self.body > 9999999] = -1

    def _com(self, a):
        return a[self.typeid == 0]

    def get_central_pos(self):
        return self._com(self.positions)

    def get_COM_velocities(self):
        return self._com(self.velocities)

    def getMass(self):
        return self._com(self.mass)

    def getOrientations(self):
        return self._com(self.orientations)

    def getAngmom(self):
        return self._com(self.angmom)

    def getMoi(self):
        return self._com(self.moment_inertia)

    def getCharges(self):
        return np.copy(self.charges)

    def getAllbody(self):
        return self.body

    def getLx(self):
        return self.Lx
