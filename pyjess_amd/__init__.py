"""MI355X-native Jess template matching with the API of pyjess.

``Jess(templates).query(molecule, rmsd_threshold, distance_cutoff,
max_dynamic_distance, ...)`` yields `Hit` objects exactly as pyjess does; the
search (candidate sets, Join enumeration, superposition) runs in hand-written
HIP kernels for gfx950 behind the C ABI of ``include/jess_amd.h``.
"""

from ._model import Atom, Mat4, Molecule, Template, TemplateAtom
from ._jess import Hit, Jess, Query

__version__ = "0.1.0"

__all__ = ["Atom", "Hit", "Jess", "Mat4", "Molecule", "Query", "Template", "TemplateAtom"]
