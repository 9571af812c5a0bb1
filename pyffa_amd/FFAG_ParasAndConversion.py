"""Constants and state conversions of the tracking path.

Mirrors the interface of the reference's ``src/FFAG_ParasAndConversion.py``:
``FFAG_GlobalParameters`` (:4-81) carries the field objects, aperture and the
physical constants into the tracker; ``FFAG_ConversionTools`` (:84-333)
converts between the (r, pr, z, pz, fi, Ek) input layout and the
(r, vr, z, vz, fi, dfi/dt) Boris / (..., E_total) RK4 state layouts.
"""
from __future__ import annotations

import numpy as np


class FFAG_GlobalParameters:
    """Container handed to the tracker (src/FFAG_ParasAndConversion.py:4-81)."""

    def __init__(self):
        self.Bmap = None
        self.Emap = None
        self.Bunch = None
        self.SEOInfo = None
        self.BHarmonics = None
        self.q = 1.60217662e-19
        self.c = 2.99792458e8
        self.E0 = 938.2723e6 * self.q
        self.m0 = self.E0 / self.c ** 2
        self.Aperture_m = None
        self.Aperture_enable = False
        self.ReverseTrack = False
        self.DumpInfo = None

    def AddBMap(self, BFieldMapData):
        self.Bmap = BFieldMapData

    def AddBHarmonics(self, BHarmonicsList):
        self.BHarmonics = BHarmonicsList

    def AddEMap(self, EFieldMapData):
        self.Emap = EFieldMapData

    def AddBunch(self, Bunch):
        self.Bunch = Bunch

    def AddSEOInfo(self, SEOdata):
        self.SEOInfo = SEOdata

    def AddDumpInfo(self, Dumpdata):
        self.DumpInfo = Dumpdata

    def SetReverseTrack(self, ReverseTrackFlag):
        self.ReverseTrack = ReverseTrackFlag

    def AddAperture(self, rmin, rmax, zmin, zmax):
        self.Aperture_m = np.array((rmin, rmax, zmin, zmax), dtype=np.float64)
        self.Aperture_enable = True


class FFAG_ConversionTools:
    """Unit conversions (src/FFAG_ParasAndConversion.py:84-333)."""

    def __init__(self):
        self.CommonVariables = FFAG_GlobalParameters()

    def _qcE0(self):
        cv = self.CommonVariables
        return cv.q, cv.c, cv.E0

    def Ek2P(self, Ek_value):
        q, c, E0 = self._qcE0()
        ek = np.asarray(Ek_value) * 1e6 * q
        return 1 / c * np.sqrt((ek + E0) ** 2 - E0 ** 2)

    def Ek2v(self, Ek_value):
        q, c, E0 = self._qcE0()
        gamma = (np.asarray(Ek_value) * 1e6 * q + E0) / E0
        return np.sqrt(1 - (1 / gamma) ** 2) * c

    def EkMeV2EtotJ(self, Ek_value):
        q, _, E0 = self._qcE0()
        return np.asarray(Ek_value) * 1e6 * q + E0

    def EtotJ2EkMeV(self, Etot_J):
        q, _, E0 = self._qcE0()
        return (np.asarray(Etot_J) - E0) / 1e6 / q

    def v2Ek(self, v):
        q, c, E0 = self._qcE0()
        gamma = 1 / np.sqrt(1 - (np.asarray(v) / c) ** 2)
        return (gamma - 1) * E0 / 1e6 / q

    def _directions(self, pr, pz):
        # pr, pz are angles: tan(pr) = vr/vfi, tan(pz) = vz/vfi
        vr_rel = np.tan(pr)
        vz_rel = np.tan(pz)
        vt = np.sqrt(vz_rel ** 2 + vr_rel ** 2 + 1.0)
        return vr_rel / vt, vz_rel / vt, 1.0 / vt

    def ConvertPrzek2Vrzek_boris(self, all_points):
        """(r, pr, z, pz, fi, Ek, ...) -> (r, vr, z, vz, fi, dfi/dt, ...)."""
        a = np.array(all_points, dtype=np.float64, copy=True)
        v = self.Ek2v(a[:, 5])
        dr, dz, df = self._directions(a[:, 1], a[:, 3])
        a[:, 1] = v * dr
        a[:, 3] = v * dz
        a[:, 5] = v * df / a[:, 0]
        return a

    def ConvertPrzek2Vrzek(self, all_points):
        """(r, pr, z, pz, fi, Ek, ...) -> (r, vr, z, vz, fi, E_total [J], ...)."""
        a = np.array(all_points, dtype=np.float64, copy=True)
        v = self.Ek2v(a[:, 5])
        dr, dz, _ = self._directions(a[:, 1], a[:, 3])
        et = self.EkMeV2EtotJ(a[:, 5])
        a[:, 1] = v * dr
        a[:, 3] = v * dz
        a[:, 5] = et
        return a

    def ConvertVrzek2Przek_boris(self, all_points):
        a = np.array(all_points, dtype=np.float64, copy=True)
        vf = a[:, 5] * a[:, 0]
        v = np.sqrt(a[:, 1] ** 2 + a[:, 3] ** 2 + vf ** 2)
        ek = self.v2Ek(v)
        dr, dz = a[:, 1] / v, a[:, 3] / v
        df = np.sqrt(1 - dr ** 2 - dz ** 2)
        a[:, 1] = np.arctan2(dr, df)
        a[:, 3] = np.arctan2(dz, df)
        a[:, 5] = ek
        return a

    def ConvertVrzek2Przek(self, all_points):
        a = np.array(all_points, dtype=np.float64, copy=True)
        ek = self.EtotJ2EkMeV(a[:, 5])
        v = self.Ek2v(ek)
        dr, dz = a[:, 1] / v, a[:, 3] / v
        df = np.sqrt(1 - dr ** 2 - dz ** 2)
        a[:, 1] = np.arctan2(dr, df)
        a[:, 3] = np.arctan2(dz, df)
        a[:, 5] = ek
        return a
