"""
Hartree atomic units and the conversion factors used at the I/O boundary.

The values are CODATA 2018, the same constants the reference package uses
(``aimd/units.py:15-26``): positions handed to a force backend are in bohr,
energies in hartree, gradients in hartree/bohr.
"""

BOHR_TO_ANG = 0.529177210903
ANG_TO_BOHR = 1.0 / BOHR_TO_ANG
AMU_TO_AU = 1822.888486209
AU_TIME_TO_FS = 2.4188843265857e-2
FS_TO_AU_TIME = 1.0 / AU_TIME_TO_FS
KB_AU = 3.166811563455546e-6
HARTREE_TO_EV = 27.211386245988
HARTREE_TO_KCALMOL = 627.5094740631
