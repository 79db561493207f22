"""
Element symbols, atomic numbers and standard atomic masses (amu) for H-Kr.

The mass table is the IUPAC standard-atomic-weight table that the reference
uses for its dynamics (``aimd/elements.py:14-28``); it has to be the same table
so that velocities drawn for the same seed and temperature agree.
"""

from __future__ import annotations

_SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar "
            "K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr").split()
_MASSES = (1.008, 4.002602, 6.94, 9.0121831, 10.81, 12.011, 14.007, 15.999,
           18.998403163, 20.1797, 22.98976928, 24.305, 26.9815385, 28.085,
           30.973761998, 32.06, 35.45, 39.948, 39.0983, 40.078, 44.955908,
           47.867, 50.9415, 51.9961, 54.938044, 55.845, 58.933194, 58.6934,
           63.546, 65.38, 69.723, 72.630, 74.921595, 78.971, 79.904, 83.798)

ATOMIC_NUMBERS: dict[str, int] = {s: z for z, s in enumerate(_SYMBOLS, start=1)}
ATOMIC_MASSES_AMU: dict[str, float] = dict(zip(_SYMBOLS, _MASSES))


def normalize_symbol(symbol: str) -> str:
    """Canonical capitalisation of an element symbol ('cl' -> 'Cl'); ValueError if unknown."""
    s = str(symbol).strip()
    sym = s[:1].upper() + s[1:].lower()
    if sym not in ATOMIC_NUMBERS:
        raise ValueError(f"Unknown element symbol: {symbol!r}")
    return sym
