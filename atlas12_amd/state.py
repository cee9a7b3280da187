"""Atmosphere state consumed by the SYNTHE hot loops.

The fields are the reference's module arrays in their Fortran layout
(column-major; a Fortran ``A(n1, n2)`` is a C-contiguous numpy array of
shape ``(n2, n1)``):

* ``wledge(nedge)``, ``continall/contabs/contscat(nf, nrhox)`` -- continuum
  edges and log10 opacities at the 3 interpolation nodes per edge interval
  (src/synthe_module.f90:250-253, :373-375);
* ``xnfpel/dopple(6*mw, nrhox)`` -- ion populations and Doppler widths
  (src/synthe_module.f90:381-382);
* ``xf_*`` depth vectors (src/synthe_module.f90:363-367), the ATLAS ``HKT``
  and ``RHOX`` (src/atlas12_modules.f90:1719, :171).

Fixture files keep only the populations of the ion slots a line list uses;
:func:`SyntheState.load` rebuilds the full ``6*mw`` layout with zeros.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MW = 139
MW6 = 6 * MW


@dataclass
class SyntheState:
    nrhox: int
    length: int
    ixwlbeg: int
    wbegin: float
    ratio: float
    resolu: float
    wledge: np.ndarray
    continall: np.ndarray  # (nrhox, nf)
    contabs: np.ndarray
    contscat: np.ndarray
    xnfpel: np.ndarray     # (nrhox, mw6)
    dopple: np.ndarray
    t: np.ndarray
    xne: np.ndarray
    rho: np.ndarray
    xnfh: np.ndarray
    xnfhe1: np.ndarray
    xnfh2: np.ndarray
    hckt: np.ndarray
    hkt_syn: np.ndarray
    hkt_atl: np.ndarray
    rhox: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def nedge(self) -> int:
        return len(self.wledge)

    @property
    def nf(self) -> int:
        return self.continall.shape[1]

    @staticmethod
    def load(path: str) -> "SyntheState":
        z = np.load(path)
        n = int(z["nrhox"])
        xp = np.zeros((n, MW6))
        dp = np.zeros((n, MW6))
        idx = z["ion_idx"]
        xp[:, idx] = z["xnfpel_sub"]
        dp[:, idx] = z["dopple_sub"]
        c = lambda k: np.ascontiguousarray(z[k], dtype=np.float64)  # noqa: E731
        return SyntheState(
            nrhox=n, length=int(z["length"]), ixwlbeg=int(z["ixwlbeg"]),
            wbegin=float(z["wbegin"]), ratio=float(z["ratio"]), resolu=float(z["resolu"]),
            wledge=c("wledge"), continall=c("continall"), contabs=c("contabs"),
            contscat=c("contscat"), xnfpel=xp, dopple=dp, t=c("t"), xne=c("xne"),
            rho=c("rho"), xnfh=c("xnfh"), xnfhe1=c("xnfhe1"), xnfh2=c("xnfh2"),
            hckt=c("hckt"), hkt_syn=c("hkt_syn"), hkt_atl=c("hkt_atl"), rhox=c("rhox"))

    def with_length(self, length: int) -> "SyntheState":
        """Same state on the first ``length`` points of the grid."""
        import copy
        s = copy.copy(self)
        s.length = int(length)
        return s


    def window(self, offset: int, length: int) -> "SyntheState":
        """The same atmosphere on another window of the same log-lambda
        grid: first point ixwlbeg + offset, ``length`` points, wbegin as
        the reference forms it (EXP(ixwlbeg * ratiolg),
        src/mod_mklinelist.f90:535-540)."""
        import copy
        s = copy.copy(self)
        s.ixwlbeg = self.ixwlbeg + int(offset)
        s.length = int(length)
        s.wbegin = float(np.exp(s.ixwlbeg * np.log(self.ratio)))
        return s

    def depths(self, idx) -> "SyntheState":
        """The state on a subset of its layers (idx ascending)."""
        import copy
        idx = np.asarray(idx)
        s = copy.copy(self)
        s.nrhox = len(idx)
        for k in ("continall", "contabs", "contscat", "xnfpel", "dopple"):
            setattr(s, k, np.ascontiguousarray(getattr(self, k)[idx]))
        for k in ("t", "xne", "rho", "xnfh", "xnfhe1", "xnfh2", "hckt", "hkt_syn", "hkt_atl",
                  "rhox"):
            setattr(s, k, np.ascontiguousarray(getattr(self, k)[idx]))
        return s


def load_coef(path: str):
    z = np.load(path)
    return (np.ascontiguousarray(z["coefj"], dtype=np.float64),
            np.ascontiguousarray(z["coefh"], dtype=np.float64))
